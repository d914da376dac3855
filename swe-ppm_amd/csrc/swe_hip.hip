// swe_hip.hip -- C ABI of the MI355X SWE block library: handle management,
// host<->device transfer and the step entry points (include/swe_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/swe_hip.h"
#include "../../include/swe_hip_mlp.h"
#include "kernels_s.hpp"
#include "kernels_u.hpp"
#include "swe_internal.h"

using namespace swe;

namespace swe {

thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return SWE_HIP_EDEVICE;
}

}  // namespace swe

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

struct swe_hip_block {
    int nx = 0, ny = 0, pitch = 0;
    float dx = 0, dy = 0, ox = 0, oy = 0;
    int solver = SOLVER_AUGRIE;
    int flavour = FLAVOUR_U;
    int strategy = 2;  // device loop: 0 recompute, 1 store + apply, 2 fused (apply on load)
    int math = 0;      // 0 IEEE (bit-compatible), 1 fast
    int xseg = 8;      // columns per wavefront strip (default_xseg)
    bool xsegSet = false;  // set by the caller: kept when a network is attached or detached
    int cur = 0;       // index of the current state buffer
    BoundaryInfo bi{{PASSIVE, PASSIVE, PASSIVE, PASSIVE}};
    float* h[2] = {nullptr, nullptr};
    float* hu[2] = {nullptr, nullptr};
    float* hv[2] = {nullptr, nullptr};
    float* b = nullptr;
    float* sums[4] = {nullptr, nullptr, nullptr, nullptr};
    float* sums2[4] = {nullptr, nullptr, nullptr, nullptr};  // second set (fused strategy)
    DeviceCtl* ctl = nullptr;
    float* dtlog = nullptr;
    int dtlogCap = 0;
    hipStream_t stream = nullptr;
    float fluxDt = -1.f;  // dt of the last flavour-S flux pass (its Q* depends on it)
    const unsigned* mlp = nullptr;  // attached learned middle state (weight fragments), or null
    swe_hip_mlp* net = nullptr;     // the attached network's handle
    uint2* deferQ = nullptr;        // deferred-edge queue of the learned middle state's fused loop
    unsigned deferCap = 0;          // its capacity: every edge of the block
    // quiet-tile record of the fused loop (kernels_fused.hpp): two slots of
    // one flag and one speed per strip tile, sized for the narrowest width
    unsigned* quiet = nullptr;
    float* tspeed = nullptr;
    unsigned* tlist = nullptr;   // tiles listed for the next fused launch, per plan group
    unsigned* tcount = nullptr;  // their number per plan group
    bool skipQuiet = true;  // skip quiet tiles (swe_hip_set_quiet_skip)
    // per-call API (swe_hip_fluxes / swe_hip_update) of flavour U: the update
    // of swe_hip_update is kept pending and applied by the next flux pass on
    // load (the fused kernel), unless another call needs the state first
    bool pending = false;    // an update of dt pendingDt is not yet applied
    bool pendingBc = false;  // swe_hip_apply_bc was called after it
    float pendingDt = 0.f;
    int pcSums = 0;  // sums set holding the last flux pass's sums (0: sums, 1: sums2)
    bool fluxValid = false;  // those sums belong to the current state (swe_hip_update may use them)
    int pcSlot = 0;  // CFL slot of the next flux pass (0, 1, 2 in turn)
    int pcStep = 0;  // fused flux passes since the last plain one: the quiet record's step
    // look-ahead (swe_hip_set_lookahead): once a caller has gone through
    // update(the dt returned) -> apply_bc -> fluxes, each flux pass queues
    // the next one at once, with the dt it returns, and that launch is kept
    // if the caller goes on in the same order
    bool lookahead = true;
    bool spec = false;     // a look-ahead launch is queued
    int specOut = 0;       // its CFL slot
    float lastDt = 0.f;    // the dt the last flux pass returned (and the look-ahead assumed)
    bool pcClean = false;  // no call but swe_hip_update / swe_hip_apply_bc since that pass
    // each flux pass's CFL maximum is written by a small kernel into a ring
    // of coherent host words, tagged with the pass's sequence number
    unsigned long long* rbRing = nullptr;
    // events of swe_hip_copy_layer into this block, made once and reused (an
    // event pair per call cost more than the copy itself)
    hipEvent_t copyEv[2] = {nullptr, nullptr};
    unsigned rbSeq = 0;  // sequence number of the last report asked for (never 0)
    // asynchronous snapshots (swe_hip_snapshot): per slot a device copy of
    // h, hu, hv taken on the block's stream and pinned host arrays filled
    // from it on snapStream while the steps go on
    struct Snap {
        float* dev[3] = {nullptr, nullptr, nullptr};
        float* host[3] = {nullptr, nullptr, nullptr};
        hipEvent_t staged = nullptr, done = nullptr;
        bool busy = false;  // queued and not yet released
    } snap[SWE_HIP_SNAPSHOT_SLOTS];
    int snapNext = 0;
    hipStream_t snapStream = nullptr;
    std::mutex snapMu;  // guards busy; never held while waiting on the GPU
    std::condition_variable snapCv;
    // Calls on one block are serialised: besides its own calls, a block is
    // read by its neighbours' layer copies, which apply its pending update
    // (drivers run blocks on several host threads).  Recursive: some entry
    // points call others.
    std::recursive_mutex mu;
};

namespace swe {
thread_local LaunchTimer g_fusedTimer;
}

namespace {

size_t field_elems(const swe_hip_block* k) { return (size_t)(k->nx + 2) * k->pitch; }

// Default strip width of a flavour-U block.  A wider segment loads and solves
// fewer halo columns per owned column and pays a wave's fixed start fewer
// times, but gives fewer waves; narrow segments spread the work more evenly
// over the SIMDs where its cost varies from edge to edge (AugRie's Newton
// iteration).  The block takes the widest of 16 and 12 whose busiest SIMD
// still holds at least the solver's threshold of waves (waves per SIMD =
// ceil(ny/62) * ceil(nx/xs) / (CUs x 4)), else 8:
//                                      16     12
//   AugRie (5 waves per SIMD)          12      4
//   learned middle state (deferring)    5      3   lighter, even waves
//   HLLE, f-wave, Hybrid                2      1   light edges
// Widths are whole numbers of the quiet record's 4-column tiles, so that
// quiet tiles can be skipped (10 columns, the old third choice, measured
// within 1% of 8 where it won).  4, 10 and 32 stay available through
// swe_hip_set_tuning.  Measured without skipping, radial dam break over
// bathymetry, us per step at 8/10/12/16/32 columns:
//   AugRie 1024^2  73.5 / 72.7 / 80.2 / 99.9 / 176   (picks 10)
//          2048^2   156 /  145 /  143 /  161 / 215   (12)
//          3000^2   299 /  266 /  259 /  268 / 309   (12)
//          4096^2   514 /  423 /  413 /  414 / 467   (16; the trench 4096^2: 424 at 12, 416 at 16)
//          8192^2  1879 / 1417 / 1387 / 1346 / 1360  (16)
//   learned 2048^2  181 /  174 /  166 /  172 / 220   (12); 3000^2: 285 at 12, 280 at 16 (16)
//   HLLE 2048^2     129 /  116 /  105 /   95 /  90   (16); f-wave 134 / 122 / 104 / 96 / 100 (16)
int default_xseg(int nx, int ny, int solver) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 8;
    const double strips = (double)((ny + kStripRows - 1) / kStripRows);
    const double simds = 4.0 * cus;
    static const double heavy[3] = {12.0, 4.0, 0.0}, learned[3] = {5.0, 3.0, 0.0}, light[3] = {2.0, 1.0, 0.0};
    const double* t = (solver == SOLVER_AUGRIE || solver == SOLVER_AUGRIE_MLP) ? heavy
                      : solver == SOLVER_AUGRIE_DEFER                         ? learned
                                                                              : light;
    // 10 columns is not a whole number of quiet-record tiles: 8 takes its place
    const int widths[3] = {16, 12, 8};
    for (int i = 0; i < 3; i++)
        if (strips * ((nx + widths[i] - 1) / widths[i]) / simds >= t[i]) return widths[i];
    return 8;
}

// AugRie with a network attached runs the learned-middle-state instantiation
int eff_solver(const swe_hip_block* k) {
    return (k->mlp && k->solver == SOLVER_AUGRIE) ? SOLVER_AUGRIE_MLP : k->solver;
}

StateView view(swe_hip_block* k, int buf) {
    StateView v;
    v.h = k->h[buf];
    v.hu = k->hu[buf];
    v.hv = k->hv[buf];
    v.b = k->b;
    v.nx = k->nx;
    v.ny = k->ny;
    v.pitch = k->pitch;
    v.dx = k->dx;
    v.dy = k->dy;
    return v;
}

SumsView sums_view(swe_hip_block* k) { return SumsView{k->sums[0], k->sums[1], k->sums[2], k->sums[3]}; }
SumsView sums2_view(swe_hip_block* k) { return SumsView{k->sums2[0], k->sums2[1], k->sums2[2], k->sums2[3]}; }
SumsView pc_sums(swe_hip_block* k) { return k->pcSums ? sums2_view(k) : sums_view(k); }

// Apply a pending per-call update now, as swe_hip_update did it eagerly
// (interior update with the dry clamp), then the boundary conditions if they
// were asked for since.
int flush_pending(swe_hip_block* k) {
    if (!k->pending) return 0;
    k->pending = false;
    k->pcStep = 0;  // updated in place: the state buffers no longer alternate
    if (int r = launch_apply_u(k->math, view(k, k->cur), pc_sums(k), k->bi, k->ctl, Slots{0, 1}, k->pendingDt, 0, nullptr,
                               0, k->stream))
        return r;
    if (k->pendingBc) {
        k->pendingBc = false;
        if (int r = launch_apply_bc(view(k, k->cur), k->bi, k->stream)) return r;
    }
    return 0;
}

// Give up a queued look-ahead launch: what it writes (the other state
// buffer and sums set, a CFL slot, a quiet record slot) is never used, and
// the quiet record starts again.
void drop_lookahead(swe_hip_block* k) {
    if (!k->spec) return;
    k->spec = false;
    k->pcStep = 0;
}

// Entry of every call other than swe_hip_update, swe_hip_apply_bc and
// swe_hip_fluxes: the state is brought to what the caller has asked for.
int settle(swe_hip_block* k) {
    drop_lookahead(k);
    k->pcClean = false;
    return flush_pending(k);
}

// ... and of calls that write the state, ghosts or boundary types (or may,
// through device pointers): the quiet record starts again.
int touch(swe_hip_block* k) {
    const int r = settle(k);
    k->pcStep = 0;
    return r;
}

// The state is replaced: nothing pending, no flux pass to update with.
void discard_per_call(swe_hip_block* k) {
    drop_lookahead(k);
    k->pending = k->pendingBc = k->fluxValid = k->pcClean = false;
    k->pcStep = 0;
}

int check(swe_hip_block* k) {
    if (!k) {
        set_error("null handle");
        return SWE_HIP_EINVAL;
    }
    return 0;
}

int ensure_second_buffer(swe_hip_block* k) {
    const int o = k->cur ^ 1;
    if (k->h[o]) return 0;
    const size_t bytes = field_elems(k) * sizeof(float);
    HIP_TRY(hipMalloc(&k->h[o], bytes));
    HIP_TRY(hipMalloc(&k->hu[o], bytes));
    HIP_TRY(hipMalloc(&k->hv[o], bytes));
    return 0;
}

int ensure_second_sums(swe_hip_block* k) {
    if (k->sums2[0]) return 0;
    const size_t bytes = field_elems(k) * sizeof(float);
    for (int i = 0; i < 4; ++i) {
        HIP_TRY(hipMalloc(&k->sums2[i], bytes));
        HIP_TRY(hipMemsetAsync(k->sums2[i], 0, bytes, k->stream));
    }
    return 0;
}

// One entry per edge the fused sweep solves: (nx+1) x-edges on each of ny
// rows and (ny+1) y-edges on each of nx columns.
int ensure_defer_queue(swe_hip_block* k) {
    if (k->deferQ) return 0;
    const size_t n = (size_t)(k->nx + 1) * k->ny + (size_t)k->nx * (k->ny + 1);
    if (n > 0xffffffffull) {
        set_error("the learned middle state's fused loop supports blocks of fewer than 2^32 edges");
        return SWE_HIP_EINVAL;
    }
    HIP_TRY(hipMalloc(&k->deferQ, n * sizeof(uint2)));
    k->deferCap = (unsigned)n;
    return 0;
}

// Strip tiles of the narrowest width: the most any width needs.
size_t max_tiles(const swe_hip_block* k) {
    return (size_t)((k->ny + kStripRows - 1) / kStripRows) * (size_t)((k->nx + 3) / 4);
}

// Workgroups the quiet-tile plan can use at any width of whole record tiles.
size_t plan_groups_max(const swe_hip_block* k) {
    const int nyt = (k->ny + kStripRows - 1) / kStripRows;
    size_t g = 0;
    for (int xs : {4, 8, 12, 16, 32}) g = std::max(g, (size_t)plan_groups(nyt, (k->nx + xs - 1) / xs, xs / kQuietCols));
    return g;
}

int ensure_quiet_record(swe_hip_block* k) {
    if (k->quiet) return 0;
    const size_t n = 2 * max_tiles(k);
    HIP_TRY(hipMalloc(&k->quiet, n * sizeof(unsigned)));
    HIP_TRY(hipMalloc(&k->tspeed, n * sizeof(float)));
    const size_t groups = plan_groups_max(k);
    HIP_TRY(hipMalloc(&k->tlist, groups * kPlanGroup * kPlanPieces * sizeof(unsigned)));
    HIP_TRY(hipMalloc(&k->tcount, groups * sizeof(unsigned)));
    return 0;
}

// A fused flux pass of the per-call API (kernels_fused.hpp): it applies the
// pending update -- of dt dtFixed, or with dtFixed < 0 of the dt that CFL
// slot slotIn gives, as a look-ahead launch does -- and the boundary
// conditions as it loads the cells, writes that state to the other buffer
// and its sums to the other set, its CFL maximum to slot `out`, and clears
// the slot after it.  The caller switches buffers and sums sets when it
// keeps the pass.  The augmented solver records quiet tiles from the first
// fused pass after a plain one and skips them from the third, as the device
// loop does: the passes between alternate the buffers with nothing else
// writing the state (every call that could resets pcStep).  rbSeq != 0:
// slot slotIn (the pass before, complete by then) is reported to the host
// under that sequence number first, by the quiet-tile plan where there is
// one, else by a readback kernel.
int pc_fused(swe_hip_block* k, int slotIn, int out, float dtFixed, unsigned rbSeq) {
    if (int r = ensure_second_buffer(k)) return r;
    if (int r = ensure_second_sums(k)) return r;
    const int o = k->cur ^ 1;
    FusedIO io{k->h[k->cur], k->hu[k->cur], k->hv[k->cur], k->b, k->h[o], k->hu[o], k->hv[o], pc_sums(k),
               k->pcSums ? sums_view(k) : sums2_view(k)};
    io.dtFixed = dtFixed;
    const int step = k->pcStep + 1;
    const int es = eff_solver(k);
    if (k->skipQuiet && es == SOLVER_AUGRIE && k->xseg % kQuietCols == 0) {
        if (int r = ensure_quiet_record(k)) return r;
        io.quiet = k->quiet;
        io.speed = k->tspeed;
        io.step = step;
        if (step >= 3) {
            if (int r = launch_quiet_plan(k->quiet, k->tspeed, k->nx, k->ny, k->xseg, step, k->ctl, slotIn, out,
                                          k->tlist, k->tcount, rbSeq, k->stream))
                return r;
            rbSeq = 0u;
            io.list = k->tlist;
            io.count = k->tcount;
        }
    }
    if (rbSeq != 0u)
        if (int r = launch_readback(k->ctl, slotIn, rbSeq, k->stream)) return r;
    if (int r = launch_fused_u(es, k->math, k->xseg, view(k, k->cur), io, k->bi, k->ctl, slotIn, out, (out + 1) % 3,
                               nullptr, 0, k->stream))
        return r;
    k->pcStep = step;
    return 0;
}

// Wait for the report of flux pass `seq` (launch_readback) and return its
// CFL maximum's bits.  The host polls the coherent ring word.  Only after
// kStreamCheckMs without it does it ask whether the stream still has work
// (so that a failed stream, or one that has run everything without the word
// arriving, ends the wait with an error): a stream query queues a marker,
// which costs the stream a pause before its next kernel.
constexpr double kStreamCheckMs = 20.0;

int wait_readback(swe_hip_block* k, unsigned seq, unsigned* bits) {
    unsigned long long* word = &k->rbRing[seq % kReadbackRing];
    auto last = std::chrono::steady_clock::now();
    for (unsigned polls = 1;; ++polls) {
        const unsigned long long v = __atomic_load_n(word, __ATOMIC_ACQUIRE);
        if ((unsigned)(v >> 32) == seq) {
            *bits = (unsigned)v;
            return 0;
        }
        if ((polls & 4095u) == 0u) {
            const auto now = std::chrono::steady_clock::now();
            if (std::chrono::duration<double, std::milli>(now - last).count() >= kStreamCheckMs) {
                last = now;
                const hipError_t q = hipStreamQuery(k->stream);
                if (q == hipSuccess) {
                    const unsigned long long w = __atomic_load_n(word, __ATOMIC_ACQUIRE);
                    if ((unsigned)(w >> 32) == seq) {
                        *bits = (unsigned)w;
                        return 0;
                    }
                    set_error("swe_hip_fluxes: the flux pass ended without its CFL maximum reaching the host");
                    return SWE_HIP_EDEVICE;
                }
                if (q != hipErrorNotReady) return hip_fail(q, "swe_hip_fluxes");
            }
        }
        __builtin_ia32_pause();
    }
}

bool same_bits(float a, float b) {
    unsigned x, y;
    std::memcpy(&x, &a, sizeof(x));
    std::memcpy(&y, &b, sizeof(y));
    return x == y;
}

// Workgroups of the deferred-edge launch: one round of waves at the learned
// solver's three waves per SIMD.
int defer_groups() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus * 3;
}

int ensure_dtlog(swe_hip_block* k, int n) {
    if (n <= k->dtlogCap) return 0;
    if (k->dtlog) HIP_TRY(hipFree(k->dtlog));
    k->dtlog = nullptr;
    HIP_TRY(hipMalloc(&k->dtlog, sizeof(float) * n));
    k->dtlogCap = n;
    return 0;
}

// Optional HIP-event instrumentation of the device loop.  Event i is
// recorded before launch i; cls[i] is the class of the work between event i
// and i+1 (0 boundary or quiet-tile plan, 1 first pass, 2 second pass).
// With the fused strategy (exact solvers) the fused launches are timed on
// their own dispatches instead (g_fusedTimer), so that no marker sits
// between the steps; their time then moves from class 0 to class 1.
struct LoopTiming {
    bool perKernel = false;
    std::vector<hipEvent_t> events;
    std::vector<int> cls;
    int used = 0;
    std::vector<hipEvent_t> fusedStart, fusedStop;  // per fused launch
    int fusedUsed = 0;
    int plans = 0;  // quiet-tile plan launches
    hipError_t mark(hipStream_t s, int nextClass) {
        if (used >= (int)events.size()) return hipSuccess;
        hipError_t e = hipEventRecord(events[used], s);
        cls.push_back(nextClass);
        ++used;
        return e;
    }
};

// Global-time-stepping loop of flavour U on the device: dt is reduced and
// applied on the GPU, the next step's boundary conditions are fused into the
// update (except after the last step, so ghosts end as in the reference).
// hostIO = false (the entry point that allocates nothing): no host transfer
// of the simulated time and no synchronisation, so the loop only queues
// kernels on the stream; dt goes to the device log k->dtlog.
int run_u_device_loop(swe_hip_block* k, int nsteps, float* t_inout, float* dt_log, LoopTiming* tm,
                      bool hostIO = true) {
    k->fluxValid = false;  // the loop's sums are not a flux pass of the per-call API
    if (int r = ensure_dtlog(k, nsteps)) return r;
    if (k->strategy == 0 || k->strategy == 2)
        if (int r = ensure_second_buffer(k)) return r;
    if (k->strategy == 2) {
        if (int r = ensure_second_sums(k)) return r;
        if (int r = ensure_quiet_record(k)) return r;
    }
    // the learned middle state's fused loop defers its hard edges to a second launch per step
    const bool defer = k->strategy == 2 && nsteps > 1 && eff_solver(k) == SOLVER_AUGRIE_MLP;
    if (defer)
        if (int r = ensure_defer_queue(k)) return r;
    const int deferGroups = defer ? defer_groups() : 0;
    float t0 = t_inout ? *t_inout : 0.f;
    if (hostIO) HIP_TRY(hipMemcpyAsync(&k->ctl->t, &t0, sizeof(float), hipMemcpyHostToDevice, k->stream));
    if (int r = launch_ctl_reset(k->ctl, 0, k->stream)) return r;
    if (tm) HIP_TRY(tm->mark(k->stream, 0));
    if (int r = launch_apply_bc(view(k, k->cur), k->bi, k->stream)) return r;
    if (k->strategy == 2) {
        // fused: sweep(STORE) on Q(0), then one launch per step that applies
        // the previous update on load, then the last update
        for (int slot = 1; slot < 3; ++slot)
            if (int r = launch_ctl_reset(k->ctl, slot, k->stream)) return r;
        SumsView set[2] = {sums_view(k), sums2_view(k)};
        if (tm && tm->perKernel) HIP_TRY(tm->mark(k->stream, 1));
        if (int r = launch_sweep_u(eff_solver(k), k->math, MODE_STORE, k->xseg, view(k, k->cur), set[0], nullptr, nullptr,
                                   nullptr, k->bi, k->ctl, Slots{0, 1}, -1.f, 0, nullptr, 0, k->stream))
            return r;
        const bool extTimed = tm && tm->perKernel && !defer;
        if (extTimed) HIP_TRY(tm->mark(k->stream, 0));  // the plans from here; fused launches timed on their own
        for (int s = 1; s < nsteps; ++s) {
            const int o = k->cur ^ 1;
            FusedIO io{k->h[k->cur], k->hu[k->cur], k->hv[k->cur], k->b, k->h[o], k->hu[o], k->hv[o],
                       set[(s - 1) & 1], set[s & 1]};
            // quiet tiles: recorded from the first fused launch, skipped from the third
            io.quiet = k->quiet;
            io.speed = k->tspeed;
            io.step = s;
            // (only the augmented solver keeps a record, at widths of whole record tiles)
            const int es = defer ? SOLVER_AUGRIE_DEFER : eff_solver(k);
            const bool record = (es == SOLVER_AUGRIE || es == SOLVER_AUGRIE_DEFER) && k->xseg % kQuietCols == 0;
            if (k->skipQuiet && s >= 3 && record) {
                if (tm && tm->perKernel && !extTimed) HIP_TRY(tm->mark(k->stream, 0));
                if (tm) ++tm->plans;
                if (int r = launch_quiet_plan(k->quiet, k->tspeed, k->nx, k->ny, k->xseg, s, k->ctl, (s + 2) % 3, s % 3,
                                              k->tlist, k->tcount, 0u, k->stream))
                    return r;
                io.list = k->tlist;
                io.count = k->tcount;
            }
            if (tm && tm->perKernel && !extTimed) HIP_TRY(tm->mark(k->stream, 1));
            if (extTimed && tm->fusedUsed < (int)tm->fusedStart.size()) {
                g_fusedTimer = LaunchTimer{tm->fusedStart[tm->fusedUsed], tm->fusedStop[tm->fusedUsed]};
                ++tm->fusedUsed;
            }
            if (defer) {
                io.deferQ = k->deferQ;
                if (int r = launch_fused_defer(k->xseg, view(k, k->cur), io, k->bi, k->ctl, (s - 1) % 3, s % 3,
                                               (s + 1) % 3, k->dtlog, s - 1, k->stream))
                    return r;
                if (int r = launch_defer_solve(view(k, o), set[s & 1], k->ctl, s % 3, k->deferQ, k->deferCap,
                                               deferGroups, k->stream))
                    return r;
            } else {
                const int r = launch_fused_u(eff_solver(k), k->math, k->xseg, view(k, k->cur), io, k->bi, k->ctl,
                                             (s - 1) % 3, s % 3, (s + 1) % 3, k->dtlog, s - 1, k->stream);
                g_fusedTimer = LaunchTimer{};
                if (r) return r;
            }
            k->cur = o;
        }
        if (tm && tm->perKernel) HIP_TRY(tm->mark(k->stream, 2));
        if (int r = launch_apply_u(k->math, view(k, k->cur), set[(nsteps - 1) & 1], k->bi, k->ctl,
                                   Slots{(nsteps - 1) % 3, nsteps % 3}, -1.f, 0, k->dtlog, nsteps - 1, k->stream))
            return r;
    }
    for (int s = 0; s < nsteps && k->strategy != 2; ++s) {
        const int par = s & 1;
        const int fuse = (s + 1 < nsteps) ? 1 : 0;  // the next step's setGhostLayer, fused
        if (tm && tm->perKernel) HIP_TRY(tm->mark(k->stream, 1));
        if (k->strategy == 0) {
            const int o = k->cur ^ 1;
            if (int r = launch_sweep_u(eff_solver(k), k->math, MODE_MAX, k->xseg, view(k, k->cur), sums_view(k), nullptr, nullptr,
                                       nullptr, k->bi, k->ctl, Slots{par, par ^ 1}, -1.f, 0, nullptr, 0, k->stream))
                return r;
            if (tm && tm->perKernel) HIP_TRY(tm->mark(k->stream, 2));
            if (int r = launch_sweep_u(eff_solver(k), k->math, MODE_UPDATE, k->xseg, view(k, k->cur), sums_view(k), k->h[o],
                                       k->hu[o], k->hv[o], k->bi, k->ctl, Slots{par, par ^ 1}, -1.f, fuse, k->dtlog, s, k->stream))
                return r;
            k->cur = o;
        } else {
            if (int r = launch_sweep_u(eff_solver(k), k->math, MODE_STORE, k->xseg, view(k, k->cur), sums_view(k), nullptr,
                                       nullptr, nullptr, k->bi, k->ctl, Slots{par, par ^ 1}, -1.f, 0, nullptr, 0, k->stream))
                return r;
            if (tm && tm->perKernel) HIP_TRY(tm->mark(k->stream, 2));
            if (int r = launch_apply_u(k->math, view(k, k->cur), sums_view(k), k->bi, k->ctl, Slots{par, par ^ 1}, -1.f, fuse,
                                       k->dtlog, s, k->stream))
                return r;
        }
    }
    if (tm) HIP_TRY(tm->mark(k->stream, 0));
    if (!hostIO) return SWE_HIP_OK;
    if (dt_log) HIP_TRY(hipMemcpyAsync(dt_log, k->dtlog, sizeof(float) * nsteps, hipMemcpyDeviceToHost, k->stream));
    float t1 = 0.f;
    HIP_TRY(hipMemcpyAsync(&t1, &k->ctl->t, sizeof(float), hipMemcpyDeviceToHost, k->stream));
    HIP_TRY(hipStreamSynchronize(k->stream));
    if (t_inout) *t_inout = t1;
    return SWE_HIP_OK;
}

// ---- file-driven scenario sampled on the device ------------------------------
// A grid as the kernel sees it: the host description plus the device copy of
// its values and the float extent the reference's strict range test uses
// (SWE_AsagiScenario.hh:100-131).
struct GridDev {
    int nx, ny;
    double x0, y0, dx, dy;
    float rx0, rx1, ry0, ry1;
    const float* v;
};

GridDev grid_dev(const swe_hip_grid& g, const float* dv) {
    GridDev d;
    d.nx = g.nx;
    d.ny = g.ny;
    d.x0 = g.x0;
    d.y0 = g.y0;
    d.dx = g.dx;
    d.dy = g.dy;
    d.rx0 = (float)(g.x0 - 0.5 * g.dx);
    d.rx1 = (float)(g.x0 + (g.nx - 0.5) * g.dx);
    d.ry0 = (float)(g.y0 - 0.5 * g.dy);
    d.ry1 = (float)(g.y0 + (g.ny - 0.5) * g.dy);
    d.v = dv;
    return d;
}

// the value of the cell containing (x, y), indices clamped to the grid
__device__ float grid_at(const GridDev& g, float x, float y) {
    long i = (long)floor(((double)x - (g.x0 - 0.5 * g.dx)) / g.dx);
    long j = (long)floor(((double)y - (g.y0 - 0.5 * g.dy)) / g.dy);
    i = i < 0 ? 0 : (i >= g.nx ? g.nx - 1 : i);
    j = j < 0 ? 0 : (j >= g.ny ? g.ny - 1 : j);
    return g.v[(size_t)j * g.nx + i];
}

// one thread per cell, ghosts included; cell (i, j) at index i * pitch + j
__global__ void grid_scenario_kernel(StateView v, float ox, float oy, GridDev bathy, GridDev displ, int hasDispl) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)(v.nx + 2) * v.pitch;
    if (c >= total) return;
    const int i = (int)(c / v.pitch), j = (int)(c % v.pitch);
    // initScenario's sample point (SWE_Block.hh:669-670): float origin + (i - 0.5) * d in double, stored as float
    const float x = (float)((double)ox + ((double)i - 0.5) * (double)v.dx);
    const float y = (float)((double)oy + ((double)j - 0.5) * (double)v.dy);
    const float bv = grid_at(bathy, x, y);
    float dv = 0.f;
    if (hasDispl && x > displ.rx0 && x < displ.rx1 && y > displ.ry0 && y < displ.ry1) dv = grid_at(displ, x, y);
    const float h = (bv > 0.f) ? 0.f : -bv;
    v.h[c] = h;
    v.hu[c] = 0.f * h;
    v.hv[c] = 0.f * h;
    const_cast<float*>(v.b)[c] = bv + dv;
}

// The analytic scenarios (SWE_simple_scenarios.hh:37-285 of the reference, as
// restated in include/swe/scenarios/SWE_simple_scenarios.hh): the same
// single-precision operations as the host classes, so the device samples
// equal the host loop's.
__device__ __forceinline__ float scen_radius(float x, float y, float cx, float cy) {
    return sqrtf((x - cx) * (x - cx) + (y - cy) * (y - cy));
}

__global__ void analytic_scenario_kernel(StateView v, float ox, float oy, int scenario) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)(v.nx + 2) * v.pitch;
    if (c >= total) return;
    const int i = (int)(c / v.pitch), j = (int)(c % v.pitch);
    const float x = (float)((double)ox + ((double)i - 0.5) * (double)v.dx);
    const float y = (float)((double)oy + ((double)j - 0.5) * (double)v.dy);
    float b = 0.f, h = 0.f;
    switch (scenario) {
        case SWE_SCENARIO_HALF_DOMAIN_DRY:
            b = -250.f;
            h = scen_radius(x, y, 250.f, 250.f) < 100.f ? 15.f : 10.f;
            break;
        case SWE_SCENARIO_RADIAL_DAM_BREAK:
            b = 10.f;
            h = scen_radius(x, y, 250.f, 250.f) < 100.f ? 15.f : 10.f;
            break;
        case SWE_SCENARIO_RADIAL_BATHYMETRY_DAM_BREAK:
            b = scen_radius(x, y, 1000.f, 1000.f) < 300.f ? 20.f : 10.f;
            h = scen_radius(x, y, 500.f, 500.f) < 300.f ? 40.f : 15.f;
            break;
        case SWE_SCENARIO_BATHYMETRY_DAM_BREAK:
            b = scen_radius(x, y, 500.f, 500.f) < 50.f ? -255.f : -260.f;
            h = 260.f;
            break;
        case SWE_SCENARIO_SEA_AT_REST: {
            const bool inner = scen_radius(x, y, 0.5f, 0.5f) < 0.1f;
            h = inner ? 9.9f : 10.0f;
            b = inner ? 0.1f : 0.0f;
            break;
        }
        case SWE_SCENARIO_SPLASHING_POOL:
            b = -250.f;
            h = 250.0f + (5.0f - (x + y) / 200.f);
            break;
        case SWE_SCENARIO_SPLASHING_CONE: {
            const float r = scen_radius(x, y, 0.5f, 0.5f);
            float hh = 4.0f - 4.5f * (r / 0.5f);
            if (r < 0.1f) hh = hh + 1.0f;
            h = (hh > 0.0f) ? hh : 0.0f;
            b = 1.0f + 9.0f * ((r < 0.5f) ? r : 0.5f);
            break;
        }
        default:  // SWE_SCENARIO_CENTRED_DAM_BREAK
            b = 0.f;
            h = scen_radius(x, y, 500.f, 500.f) < 100.f ? 15.f : 10.f;
            break;
    }
    v.h[c] = h;
    v.hu[c] = 0.f * h;
    v.hv[c] = 0.f * h;
    const_cast<float*>(v.b)[c] = b;
}

// exp of a single-precision argument, evaluated in double and rounded once
__device__ __forceinline__ float exp_rounded(float q) { return (float)exp((double)q); }

// The seeded workloads (swe_hip_seeded): the single-precision operations of
// swe_ppm_amd.scenarios.sample_seeded in the same order.
__global__ void seeded_scenario_kernel(StateView v, float ox, float oy, swe_hip_seeded p) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)(v.nx + 2) * v.pitch;
    if (c >= total) return;
    const int i = (int)(c / v.pitch), j = (int)(c % v.pitch);
    const float x = (float)((double)ox + ((double)i - 0.5) * (double)v.dx);
    const float y = (float)((double)oy + ((double)j - 0.5) * (double)v.dy);
    float b, h;
    if (p.kind == SWE_SEEDED_TRENCH) {
        const float ax = x - p.trench_x;
        const float depth = 10.f + 3990.f * exp_rounded(-(ax * ax) / p.trench_w2);
        const float hx = x - p.hump_x, hy = y - p.hump_y;
        const float eta = 2.f * exp_rounded(-(hx * hx + hy * hy) / p.hump_r2);
        b = -depth;
        h = depth + eta;
    } else {
        b = 0.f;
        for (int k = 0; k < p.nbumps; ++k) {
            const float bx = x - p.bump_x[k], by = y - p.bump_y[k];
            b = b + p.bump_a[k] * exp_rounded(-(bx * bx + by * by) / p.bump_s2[k]);
        }
        const float eta = scen_radius(x, y, p.dam_x, p.dam_y) < p.dam_radius ? p.eta_in : p.eta_out;
        const float d = eta - b;
        h = d > 0.f ? d : 0.f;
    }
    v.h[c] = h;
    v.hu[c] = 0.f;
    v.hv[c] = 0.f;
    const_cast<float*>(v.b)[c] = b;
}

bool grid_ok(const swe_hip_grid* g) {
    return g && g->values && g->nx >= 1 && g->ny >= 1 && g->dx > 0.0 && g->dy > 0.0;
}

}  // namespace

namespace swe {
int block_view(swe_hip_block* k, StateView* v, hipStream_t* stream) {
    if (!k || k->flavour != FLAVOUR_U) {
        set_error("block_view: a flavour-U block is needed");
        return SWE_HIP_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    *v = view(k, k->cur);
    *stream = k->stream;
    return 0;
}
}  // namespace swe

extern "C" {

const char* swe_hip_last_error(void) { return g_last_error; }

const char* swe_hip_version(void) { return "swe_hip 0.1 gfx950"; }

int swe_hip_create(int nx, int ny, float dx, float dy, float ox, float oy, int solver, int flavour,
                   swe_hip_block** out) {
    if (!out || nx < 1 || ny < 1 || !(dx > 0.f) || !(dy > 0.f) || solver < 0 || solver > 3 ||
        (flavour != FLAVOUR_U && flavour != FLAVOUR_S)) {
        set_error("swe_hip_create: invalid argument (nx=%d ny=%d dx=%g dy=%g solver=%d flavour=%d)", nx, ny, dx,
                  dy, solver, flavour);
        return SWE_HIP_EINVAL;
    }
    if ((size_t)(nx + 2) * (size_t)(ny + 2) > ((size_t)1 << 31)) {
        set_error("swe_hip_create: block too large");
        return SWE_HIP_EINVAL;
    }
    swe_hip_block* k = new (std::nothrow) swe_hip_block();
    if (!k) {
        set_error("swe_hip_create: out of host memory");
        return SWE_HIP_EINVAL;
    }
    k->nx = nx;
    k->ny = ny;
    k->pitch = ny + 2;
    k->dx = dx;
    k->dy = dy;
    k->ox = ox;
    k->oy = oy;
    k->solver = solver;
    k->flavour = flavour;
    k->xseg = default_xseg(nx, ny, solver);
    const size_t bytes = field_elems(k) * sizeof(float);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&k->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&k->rbRing, kReadbackRing * sizeof(unsigned long long), hipHostMallocCoherent);
    if (e == hipSuccess) std::memset(k->rbRing, 0, kReadbackRing * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&k->h[0], bytes);
    if (e == hipSuccess) e = hipMalloc(&k->hu[0], bytes);
    if (e == hipSuccess) e = hipMalloc(&k->hv[0], bytes);
    if (e == hipSuccess) e = hipMalloc(&k->b, bytes);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipMalloc(&k->sums[i], bytes);
    if (e == hipSuccess) e = hipMalloc(&k->ctl, sizeof(DeviceCtl));
    if (e == hipSuccess) e = hipMemsetAsync(k->ctl, 0, sizeof(DeviceCtl), k->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&k->ctl->ring, &k->rbRing, sizeof(k->rbRing), hipMemcpyHostToDevice, k->stream);
    for (int i = 0; i < 4 && e == hipSuccess; ++i) e = hipMemsetAsync(k->sums[i], 0, bytes, k->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->h[0], 0, bytes, k->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->hu[0], 0, bytes, k->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->hv[0], 0, bytes, k->stream);
    if (e == hipSuccess) e = hipMemsetAsync(k->b, 0, bytes, k->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(k->stream);
    if (e != hipSuccess) {
        hip_fail(e, "swe_hip_create");
        swe_hip_destroy(k);
        return SWE_HIP_EDEVICE;
    }
    *out = k;
    return SWE_HIP_OK;
}

int swe_hip_destroy(swe_hip_block* k) {
    if (!k) return SWE_HIP_OK;
    if (k->stream) (void)hipStreamSynchronize(k->stream);
    if (k->snapStream) (void)hipStreamSynchronize(k->snapStream);
    for (auto& sn : k->snap) {
        for (int i = 0; i < 3; ++i) {
            if (sn.dev[i]) (void)hipFree(sn.dev[i]);
            if (sn.host[i]) (void)hipHostFree(sn.host[i]);
        }
        if (sn.staged) (void)hipEventDestroy(sn.staged);
        if (sn.done) (void)hipEventDestroy(sn.done);
    }
    if (k->snapStream) (void)hipStreamDestroy(k->snapStream);
    if (k->net) swe::mlp_detach(k->net, k->stream);
    for (int i = 0; i < 2; ++i) {
        if (k->h[i]) (void)hipFree(k->h[i]);
        if (k->hu[i]) (void)hipFree(k->hu[i]);
        if (k->hv[i]) (void)hipFree(k->hv[i]);
    }
    if (k->b) (void)hipFree(k->b);
    for (int i = 0; i < 4; ++i) {
        if (k->sums[i]) (void)hipFree(k->sums[i]);
        if (k->sums2[i]) (void)hipFree(k->sums2[i]);
    }
    if (k->ctl) (void)hipFree(k->ctl);
    if (k->dtlog) (void)hipFree(k->dtlog);
    if (k->deferQ) (void)hipFree(k->deferQ);
    if (k->quiet) (void)hipFree(k->quiet);
    if (k->tspeed) (void)hipFree(k->tspeed);
    if (k->tlist) (void)hipFree(k->tlist);
    if (k->tcount) (void)hipFree(k->tcount);
    if (k->rbRing) (void)hipHostFree(k->rbRing);
    for (hipEvent_t ev : k->copyEv)
        if (ev) (void)hipEventDestroy(ev);
    if (k->stream) (void)hipStreamDestroy(k->stream);
    delete k;
    return SWE_HIP_OK;
}

int swe_hip_upload(swe_hip_block* k, const float* h, const float* hu, const float* hv, const float* b) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    discard_per_call(k);  // the state is replaced
    if (!h || !hu || !hv || !b) {
        set_error("swe_hip_upload: null array");
        return SWE_HIP_EINVAL;
    }
    const size_t bytes = field_elems(k) * sizeof(float);
    HIP_TRY(hipMemcpyAsync(k->h[k->cur], h, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->hu[k->cur], hu, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->hv[k->cur], hv, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipMemcpyAsync(k->b, b, bytes, hipMemcpyHostToDevice, k->stream));
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_download(swe_hip_block* k, float* h, float* hu, float* hv, float* b) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = settle(k)) return r;
    const size_t bytes = field_elems(k) * sizeof(float);
    if (h) HIP_TRY(hipMemcpyAsync(h, k->h[k->cur], bytes, hipMemcpyDeviceToHost, k->stream));
    if (hu) HIP_TRY(hipMemcpyAsync(hu, k->hu[k->cur], bytes, hipMemcpyDeviceToHost, k->stream));
    if (hv) HIP_TRY(hipMemcpyAsync(hv, k->hv[k->cur], bytes, hipMemcpyDeviceToHost, k->stream));
    if (b) HIP_TRY(hipMemcpyAsync(b, k->b, bytes, hipMemcpyDeviceToHost, k->stream));
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_set_boundary(swe_hip_block* k, int side, int type) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (side < BND_LEFT || side > BND_TOP || type < OUTFLOW || type > PASSIVE) {
        set_error("swe_hip_set_boundary: invalid side %d or type %d", side, type);
        return SWE_HIP_EINVAL;
    }
    if (type == INFLOW) {
        // the reference asserts on INFLOW in applyBoundaryConditions (SWE_Block.hh:784-786)
        set_error("swe_hip_set_boundary: INFLOW boundaries are not supported (the reference asserts on them)");
        return SWE_HIP_EINVAL;
    }
    k->bi.type[side] = type;
    return SWE_HIP_OK;
}

int swe_hip_apply_bc(swe_hip_block* k) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (k->pending) {  // the next flux pass writes these ghosts as it applies the update
        k->pendingBc = true;
        return SWE_HIP_OK;
    }
    // ghosts of the current state rewritten: not the order a look-ahead assumes
    drop_lookahead(k);
    k->pcClean = false;
    k->pcStep = 0;
    return launch_apply_bc(view(k, k->cur), k->bi, k->stream);
}

int swe_hip_apply_bc_bathymetry(swe_hip_block* k) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (int r = launch_apply_bc_bathymetry(k->b, k->nx, k->ny, k->pitch, k->bi, k->stream)) return r;
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_init_grid_scenario(swe_hip_block* k, const swe_hip_grid* bathymetry, const swe_hip_grid* displacement) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    discard_per_call(k);  // the state is replaced
    if (!grid_ok(bathymetry) || (displacement && !grid_ok(displacement))) {
        set_error("swe_hip_init_grid_scenario: a grid needs values, nx, ny >= 1 and positive spacing");
        return SWE_HIP_EINVAL;
    }
    const size_t nb = (size_t)bathymetry->nx * bathymetry->ny;
    const size_t nd = displacement ? (size_t)displacement->nx * displacement->ny : 0;
    float* dev = nullptr;
    HIP_TRY(hipMalloc(&dev, sizeof(float) * (nb + nd + 1)));
    hipError_t e = hipMemcpyAsync(dev, bathymetry->values, sizeof(float) * nb, hipMemcpyHostToDevice, k->stream);
    if (e == hipSuccess && displacement)
        e = hipMemcpyAsync(dev + nb, displacement->values, sizeof(float) * nd, hipMemcpyHostToDevice, k->stream);
    if (e == hipSuccess) {
        const GridDev gb = grid_dev(*bathymetry, dev);
        const GridDev gd = displacement ? grid_dev(*displacement, dev + nb) : gb;
        const long total = (long)field_elems(k);
        hipLaunchKernelGGL(grid_scenario_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, k->stream,
                           view(k, k->cur), k->ox, k->oy, gb, gd, displacement ? 1 : 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(k->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_fail(e, "swe_hip_init_grid_scenario");
    // applyBoundaryConditions, then applyBoundaryBathymetry (SWE_Block.hh:679-681)
    if (int r = launch_apply_bc(view(k, k->cur), k->bi, k->stream)) return r;
    if (int r = launch_apply_bc_bathymetry(k->b, k->nx, k->ny, k->pitch, k->bi, k->stream)) return r;
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_init_scenario(swe_hip_block* k, int scenario) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    discard_per_call(k);  // the state is replaced
    if (scenario < SWE_SCENARIO_HALF_DOMAIN_DRY || scenario > SWE_SCENARIO_CENTRED_DAM_BREAK) {
        set_error("swe_hip_init_scenario: unknown scenario %d", scenario);
        return SWE_HIP_EINVAL;
    }
    const long total = (long)field_elems(k);
    hipLaunchKernelGGL(analytic_scenario_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, k->stream,
                       view(k, k->cur), k->ox, k->oy, scenario);
    HIP_TRY(hipGetLastError());
    // applyBoundaryConditions, then applyBoundaryBathymetry (SWE_Block.hh:679-681)
    if (int r = launch_apply_bc(view(k, k->cur), k->bi, k->stream)) return r;
    if (int r = launch_apply_bc_bathymetry(k->b, k->nx, k->ny, k->pitch, k->bi, k->stream)) return r;
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_init_seeded(swe_hip_block* k, const swe_hip_seeded* workload) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (!workload || (workload->kind != SWE_SEEDED_DAM && workload->kind != SWE_SEEDED_TRENCH) ||
        workload->nbumps < 0 || workload->nbumps > SWE_SEEDED_MAX_BUMPS) {
        set_error("swe_hip_init_seeded: invalid workload (kind %d, %d bumps)", workload ? workload->kind : -1,
                  workload ? workload->nbumps : -1);
        return SWE_HIP_EINVAL;
    }
    discard_per_call(k);  // the state is replaced
    const long total = (long)field_elems(k);
    hipLaunchKernelGGL(seeded_scenario_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, k->stream,
                       view(k, k->cur), k->ox, k->oy, *workload);
    HIP_TRY(hipGetLastError());
    // applyBoundaryConditions, then applyBoundaryBathymetry (SWE_Block.hh:679-681)
    if (int r = launch_apply_bc(view(k, k->cur), k->bi, k->stream)) return r;
    if (int r = launch_apply_bc_bathymetry(k->b, k->nx, k->ny, k->pitch, k->bi, k->stream)) return r;
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_fluxes(swe_hip_block* k, float* max_timestep) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (!max_timestep) {
        set_error("swe_hip_fluxes: null output");
        return SWE_HIP_EINVAL;
    }
    unsigned int bits = 0;
    if (k->flavour == FLAVOUR_U) {
        // an update left pending by swe_hip_update, with the boundary
        // conditions asked for since (the drivers' order: updateUnknowns,
        // setGhostLayer, computeNumericalFluxes), is applied by the fused
        // kernel as it loads the cells -- one launch for the whole step.
        // Without the boundary conditions it is applied on its own first.
        const bool inOrder = k->pcClean && k->pending && k->pendingBc && same_bits(k->pendingDt, k->lastDt);
        int out;
        if (k->spec && inOrder) {
            // the look-ahead launch made exactly this pass: keep it
            k->spec = false;
            out = k->specOut;
            k->pending = k->pendingBc = false;
            k->cur ^= 1;
            k->pcSums ^= 1;
        } else {
            drop_lookahead(k);
            if (k->pending && !k->pendingBc)
                if (int r = flush_pending(k)) return r;
            out = k->pcSlot;
            k->pcSlot = (out + 1) % 3;
            if (int r = launch_ctl_reset(k->ctl, out, k->stream)) return r;
            if (k->pending) {
                if (int r = pc_fused(k, (out + 2) % 3, out, k->pendingDt, 0u)) return r;
                k->pending = k->pendingBc = false;
                k->cur ^= 1;
                k->pcSums ^= 1;
            } else {
                if (int r = launch_sweep_u(eff_solver(k), k->math, MODE_STORE, k->xseg, view(k, k->cur), pc_sums(k),
                                           nullptr, nullptr, nullptr, k->bi, k->ctl, Slots{out, (out + 1) % 3}, -1.f,
                                           0, nullptr, 0, k->stream))
                    return r;
                k->pcStep = 0;
            }
        }
        k->fluxValid = true;
        if (++k->rbSeq == 0u) k->rbSeq = 1u;  // 0 means "no report"
        const unsigned want = k->rbSeq;
        if (inOrder && k->lookahead && !k->mlp) {
            // queue the next pass with the dt being returned while the host
            // waits, if the caller keeps to the order just seen (the device
            // forms that dt from the same maximum by the same IEEE
            // operations); it reports this pass's maximum first
            const int next = k->pcSlot;
            k->pcSlot = (next + 1) % 3;
            if (int r = pc_fused(k, out, next, -1.f, want)) return r;
            k->spec = true;
            k->specOut = next;
        } else if (int r = launch_readback(k->ctl, out, want, k->stream)) {
            return r;
        }
        if (int r = wait_readback(k, want, &bits)) return r;
        float s;
        std::memcpy(&s, &bits, sizeof(s));
        const float dt = host_timestep_u(s, k->dx, k->dy);
        k->lastDt = dt;
        k->pcClean = true;
        *max_timestep = dt;
    } else {
        if (int r = launch_ctl_reset(k->ctl, 0, k->stream)) return r;
        if (int r = launch_sweep_s_x(k->solver, view(k, k->cur), sums_view(k), k->ctl, 0, k->stream)) return r;
        HIP_TRY(hipMemcpyAsync(&bits, &k->ctl->smax[0], sizeof(bits), hipMemcpyDeviceToHost, k->stream));
        HIP_TRY(hipStreamSynchronize(k->stream));
        float s;
        std::memcpy(&s, &bits, sizeof(s));
        const float dt = (float).4 * (k->dx / s);  // SWE_DimensionalSplitting.cpp:221, no zero guard
        if (int r = launch_sweep_s_y(k->solver, view(k, k->cur), sums_view(k), k->stream)) return r;
        k->fluxDt = dt;
        *max_timestep = dt;
    }
    return SWE_HIP_OK;
}

int swe_hip_update(swe_hip_block* k, float dt) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (k->flavour == FLAVOUR_U) {
        if (!(dt >= 0.f)) {
            set_error("swe_hip_update: dt must be >= 0");
            return SWE_HIP_EINVAL;
        }
        if (!k->fluxValid) {
            set_error("swe_hip_update: no flux pass of the current state (call swe_hip_fluxes first; device loops, "
                      "uploads and scenario initialisation discard the last one)");
            return SWE_HIP_ESTATE;
        }
        k->fluxValid = false;  // the sums are used up
        k->pending = true;
        k->pendingBc = false;
        k->pendingDt = dt;
        return SWE_HIP_OK;
    }
    // flavour S: Q* was formed with the flux pass's own dt (the reference
    // asserts that dt equals it, SWE_DimensionalSplitting.cpp:278)
    if (k->fluxDt < 0.f) {
        set_error("swe_hip_update: flavour S needs swe_hip_fluxes first");
        return SWE_HIP_ESTATE;
    }
    return launch_apply_s(view(k, k->cur), sums_view(k), k->fluxDt, k->stream);
}

int swe_hip_run(swe_hip_block* k, int nsteps, float* t_inout, float* dt_log) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (nsteps < 0) {
        set_error("swe_hip_run: nsteps < 0");
        return SWE_HIP_EINVAL;
    }
    if (nsteps == 0) return SWE_HIP_OK;
    if (k->flavour != FLAVOUR_U) {
        // flavour S keeps its host-visible dt (swe_simple reads it every step)
        float t = t_inout ? *t_inout : 0.f;
        for (int s = 0; s < nsteps; ++s) {
            float dt;
            if (int r = swe_hip_apply_bc(k)) return r;
            if (int r = swe_hip_fluxes(k, &dt)) return r;
            if (int r = swe_hip_update(k, dt)) return r;
            if (dt_log) dt_log[s] = dt;
            t += dt;
        }
        HIP_TRY(hipStreamSynchronize(k->stream));
        if (t_inout) *t_inout = t;
        return SWE_HIP_OK;
    }
    return run_u_device_loop(k, nsteps, t_inout, dt_log, nullptr);
}

int swe_hip_run_timed(swe_hip_block* k, int nsteps, float* t_inout, float* total_ms, float* kernel_ms,
                      int* kernel_launches) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (nsteps < 1 || k->flavour != FLAVOUR_U) {
        set_error("swe_hip_run_timed: needs nsteps >= 1 and flavour U");
        return SWE_HIP_EINVAL;
    }
    LoopTiming tm;
    tm.perKernel = kernel_ms != nullptr;
    const int nev = tm.perKernel ? 3 * nsteps + 3 : 2;
    tm.events.resize(nev);
    for (auto& e : tm.events) HIP_TRY(hipEventCreate(&e));
    if (tm.perKernel && k->strategy == 2) {
        tm.fusedStart.resize(nsteps);
        tm.fusedStop.resize(nsteps);
        for (auto& e : tm.fusedStart) HIP_TRY(hipEventCreate(&e));
        for (auto& e : tm.fusedStop) HIP_TRY(hipEventCreate(&e));
    }
    int rc = run_u_device_loop(k, nsteps, t_inout, nullptr, &tm);
    if (rc == 0) {
        float ms = 0.f;
        hipError_t e = hipEventElapsedTime(&ms, tm.events[0], tm.events[tm.used - 1]);
        if (e != hipSuccess) rc = hip_fail(e, "hipEventElapsedTime");
        if (total_ms) *total_ms = ms;
        if (rc == 0 && tm.perKernel) {
            float acc[3] = {0.f, 0.f, 0.f};
            int cnt[3] = {0, 0, 0};
            for (int i = 0; i + 1 < tm.used; ++i) {
                float d = 0.f;
                e = hipEventElapsedTime(&d, tm.events[i], tm.events[i + 1]);
                if (e != hipSuccess) {
                    rc = hip_fail(e, "hipEventElapsedTime");
                    break;
                }
                acc[tm.cls[i]] += d;
                cnt[tm.cls[i]] += 1;
            }
            if (tm.fusedUsed > 0) {  // the fused launches ran inside the class-0 stretch
                float fused = 0.f;
                for (int i = 0; i < tm.fusedUsed && rc == 0; ++i) {
                    float d = 0.f;
                    e = hipEventElapsedTime(&d, tm.fusedStart[i], tm.fusedStop[i]);
                    if (e != hipSuccess) rc = hip_fail(e, "hipEventElapsedTime");
                    fused += d;
                }
                acc[0] -= fused;
                acc[1] += fused;
                cnt[0] += tm.plans - 1;
                cnt[1] += tm.fusedUsed;
            }
            for (int c = 0; c < 3; ++c) {
                kernel_ms[c] = acc[c];
                if (kernel_launches) kernel_launches[c] = cnt[c];
            }
        }
    }
    for (auto& e : tm.events) (void)hipEventDestroy(e);
    for (auto& e : tm.fusedStart) (void)hipEventDestroy(e);
    for (auto& e : tm.fusedStop) (void)hipEventDestroy(e);
    return rc;
}

int swe_hip_synchronize(swe_hip_block* k) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = settle(k)) return r;
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_device_pointers(swe_hip_block* k, float** h, float** hu, float** hv, float** b, int* pitch) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (h) *h = k->h[k->cur];
    if (hu) *hu = k->hu[k->cur];
    if (hv) *hv = k->hv[k->cur];
    if (b) *b = k->b;
    if (pitch) *pitch = k->pitch;
    return SWE_HIP_OK;
}

}  // extern "C"

namespace {

// Geometry of one boundary line: element 1..n of the ghost line (ghost=1)
// or of the adjacent copy layer (ghost=0), as (offset, stride) in floats.
struct LineGeom {
    size_t off;
    size_t stride;
    int n;
};

LineGeom line_geom(const swe_hip_block* k, int side, int ghost) {
    LineGeom g;
    switch (side) {
        case BND_LEFT:
            g.off = (size_t)(ghost ? 0 : 1) * k->pitch + 1;
            g.stride = 1;
            g.n = k->ny;
            break;
        case BND_RIGHT:
            g.off = (size_t)(ghost ? k->nx + 1 : k->nx) * k->pitch + 1;
            g.stride = 1;
            g.n = k->ny;
            break;
        case BND_BOTTOM:
            g.off = (size_t)1 * k->pitch + (ghost ? 0 : 1);
            g.stride = k->pitch;
            g.n = k->nx;
            break;
        default:
            g.off = (size_t)1 * k->pitch + (ghost ? k->ny + 1 : k->ny);
            g.stride = k->pitch;
            g.n = k->nx;
            break;
    }
    return g;
}

float* field_ptr(swe_hip_block* k, int bit) {
    switch (bit) {
        case SWE_FIELD_H:
            return k->h[k->cur];
        case SWE_FIELD_HU:
            return k->hu[k->cur];
        case SWE_FIELD_HV:
            return k->hv[k->cur];
        default:
            return k->b;
    }
}

int opposite(int side) { return side ^ 1; }

}  // namespace

extern "C" {

int swe_hip_snapshot(swe_hip_block* k, int* slot) {
    if (int r = check(k)) return r;
    if (!slot) {
        set_error("swe_hip_snapshot: null slot");
        return SWE_HIP_EINVAL;
    }
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = settle(k)) return r;
    const int s = k->snapNext;
    auto& sn = k->snap[s];
    {  // double buffering: the slot's last snapshot must have been released
        std::unique_lock<std::mutex> g(k->snapMu);
        k->snapCv.wait(g, [&] { return !sn.busy; });
    }
    const size_t bytes = field_elems(k) * sizeof(float);
    if (!k->snapStream) HIP_TRY(hipStreamCreateWithFlags(&k->snapStream, hipStreamNonBlocking));
    for (int i = 0; i < 3; ++i) {
        if (!sn.dev[i]) HIP_TRY(hipMalloc(&sn.dev[i], bytes));
        if (!sn.host[i]) HIP_TRY(hipHostMalloc((void**)&sn.host[i], bytes, hipHostMallocDefault));
    }
    if (!sn.staged) HIP_TRY(hipEventCreateWithFlags(&sn.staged, hipEventDisableTiming));
    if (!sn.done) HIP_TRY(hipEventCreateWithFlags(&sn.done, hipEventDisableTiming));
    // a device copy in the block's stream order (later steps may overwrite
    // the state), then the transfer on the snapshot stream
    const float* src[3] = {k->h[k->cur], k->hu[k->cur], k->hv[k->cur]};
    for (int i = 0; i < 3; ++i)
        HIP_TRY(hipMemcpyAsync(sn.dev[i], src[i], bytes, hipMemcpyDeviceToDevice, k->stream));
    HIP_TRY(hipEventRecord(sn.staged, k->stream));
    HIP_TRY(hipStreamWaitEvent(k->snapStream, sn.staged, 0));
    for (int i = 0; i < 3; ++i)
        HIP_TRY(hipMemcpyAsync(sn.host[i], sn.dev[i], bytes, hipMemcpyDeviceToHost, k->snapStream));
    HIP_TRY(hipEventRecord(sn.done, k->snapStream));
    {
        std::lock_guard<std::mutex> g(k->snapMu);
        sn.busy = true;
    }
    k->snapNext = (s + 1) % SWE_HIP_SNAPSHOT_SLOTS;
    *slot = s;
    return SWE_HIP_OK;
}

int swe_hip_snapshot_wait(swe_hip_block* k, int slot, const float** h, const float** hu, const float** hv) {
    if (int r = check(k)) return r;
    if (slot < 0 || slot >= SWE_HIP_SNAPSHOT_SLOTS || !h || !hu || !hv) {
        set_error("swe_hip_snapshot_wait: invalid slot %d or null output", slot);
        return SWE_HIP_EINVAL;
    }
    auto& sn = k->snap[slot];
    hipEvent_t done = nullptr;
    {
        std::lock_guard<std::mutex> g(k->snapMu);
        if (!sn.busy) {
            set_error("swe_hip_snapshot_wait: slot %d holds no snapshot", slot);
            return SWE_HIP_ESTATE;
        }
        done = sn.done;
    }
    HIP_TRY(hipEventSynchronize(done));
    *h = sn.host[0];
    *hu = sn.host[1];
    *hv = sn.host[2];
    return SWE_HIP_OK;
}

int swe_hip_snapshot_release(swe_hip_block* k, int slot) {
    if (int r = check(k)) return r;
    if (slot < 0 || slot >= SWE_HIP_SNAPSHOT_SLOTS) {
        set_error("swe_hip_snapshot_release: invalid slot %d", slot);
        return SWE_HIP_EINVAL;
    }
    auto& sn = k->snap[slot];
    {
        std::lock_guard<std::mutex> g(k->snapMu);
        if (!sn.busy) {
            set_error("swe_hip_snapshot_release: slot %d holds no snapshot", slot);
            return SWE_HIP_ESTATE;
        }
    }
    // the transfer must be over before the slot is reused
    HIP_TRY(hipEventSynchronize(sn.done));
    {
        std::lock_guard<std::mutex> g(k->snapMu);
        sn.busy = false;
    }
    k->snapCv.notify_all();
    return SWE_HIP_OK;
}

int swe_hip_copy_layer(swe_hip_block* dst, int side, swe_hip_block* src, int fields) {
    if (int r = check(dst)) return r;
    if (int r = check(src)) return r;
    if (side < BND_LEFT || side > BND_TOP || fields <= 0 || fields > 15) {
        set_error("swe_hip_copy_layer: invalid side %d or fields %d", side, fields);
        return SWE_HIP_EINVAL;
    }
    std::unique_lock<std::recursive_mutex> lockDst(dst->mu, std::defer_lock), lockSrc(src->mu, std::defer_lock);
    if (dst == src)
        lockDst.lock();
    else
        std::lock(lockDst, lockSrc);
    // The copy layer must be the source's state with its latest update.  A
    // source update still pending stays pending (its next flux pass applies
    // it on load): h, hu, hv are copied with the update applied on the fly,
    // the same arithmetic as applying it.  The destination's pending update
    // keeps the ghost lines of CONNECT and PASSIVE sides as they are (the next
    // flux pass carries them over), but rewrites those of OUTFLOW and WALL
    // sides, so a copy into such a line waits for it.  A look-ahead launch of
    // the destination read the old ghost line.
    const int moving = fields & (SWE_FIELD_H | SWE_FIELD_HU | SWE_FIELD_HV);
    const bool onTheFly = src->flavour == FLAVOUR_U && src->pending && src != dst &&
                          (moving == 0 || moving == (SWE_FIELD_H | SWE_FIELD_HU | SWE_FIELD_HV));
    if (!onTheFly)
        if (int r = settle(src)) return r;
    drop_lookahead(dst);
    dst->pcClean = false;
    dst->pcStep = 0;
    if (dst->bi.type[side] == OUTFLOW || dst->bi.type[side] == WALL)
        if (int r = flush_pending(dst)) return r;
    const LineGeom gd = line_geom(dst, side, 1);
    const LineGeom gs = line_geom(src, opposite(side), 0);
    if (gd.n != gs.n) {
        set_error("swe_hip_copy_layer: edge lengths differ (%d vs %d)", gd.n, gs.n);
        return SWE_HIP_EINVAL;
    }
    for (hipEvent_t& ev : dst->copyEv)
        if (!ev) HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipEvent_t e1 = dst->copyEv[0], e2 = dst->copyEv[1];
    hipError_t e = hipEventRecord(e1, src->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(dst->stream, e1, 0);
    int rc = 0;
    if (onTheFly && moving && e == hipSuccess) {
        const SumsView ps = pc_sums(src);
        LinePending L;
        L.dh = dst->h[dst->cur] + gd.off;
        L.dhu = dst->hu[dst->cur] + gd.off;
        L.dhv = dst->hv[dst->cur] + gd.off;
        L.dstride = gd.stride;
        L.sh = src->h[src->cur] + gs.off;
        L.shu = src->hu[src->cur] + gs.off;
        L.shv = src->hv[src->cur] + gs.off;
        L.sums = SumsView{ps.sx + gs.off, ps.sy + gs.off, ps.su + gs.off, ps.sv + gs.off};
        L.sstride = gs.stride;
        L.n = gs.n;
        L.dt = src->pendingDt;
        L.dx = src->dx;
        L.dy = src->dy;
        rc = src->math == 1 ? launch_copy_line_pending_variant<1>(L, dst->stream)
                            : launch_copy_line_pending_variant<0>(L, dst->stream);
    }
    for (int bit = 1; bit <= 8 && e == hipSuccess && rc == 0; bit <<= 1) {
        if (!(fields & bit)) continue;
        if (onTheFly && bit != SWE_FIELD_B) continue;  // copied above
        float* d = field_ptr(dst, bit) + gd.off;
        const float* s = field_ptr(src, bit) + gs.off;
        e = hipMemcpy2DAsync(d, gd.stride * sizeof(float), s, gs.stride * sizeof(float), sizeof(float), gd.n,
                             hipMemcpyDeviceToDevice, dst->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(e2, dst->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(src->stream, e2, 0);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "swe_hip_copy_layer");
    return SWE_HIP_OK;
}

static int layer_host_transfer(swe_hip_block* k, int side, int ghost, int fields, float* out, const float* in) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (side < BND_LEFT || side > BND_TOP || fields <= 0 || fields > 15 || (!out && !in)) {
        set_error("layer transfer: invalid side %d, fields %d or buffer", side, fields);
        return SWE_HIP_EINVAL;
    }
    if (int r = out ? settle(k) : touch(k)) return r;
    const LineGeom g = line_geom(k, side, ghost ? 1 : 0);
    size_t pos = 0;
    for (int bit = 1; bit <= 8; bit <<= 1) {
        if (!(fields & bit)) continue;
        float* d = field_ptr(k, bit) + g.off;
        if (out)
            HIP_TRY(hipMemcpy2DAsync(out + pos, sizeof(float), d, g.stride * sizeof(float), sizeof(float), g.n,
                                     hipMemcpyDeviceToHost, k->stream));
        else
            HIP_TRY(hipMemcpy2DAsync(d, g.stride * sizeof(float), in + pos, sizeof(float), sizeof(float), g.n,
                                     hipMemcpyHostToDevice, k->stream));
        pos += g.n;
    }
    HIP_TRY(hipStreamSynchronize(k->stream));
    return SWE_HIP_OK;
}

int swe_hip_read_layer(swe_hip_block* k, int side, int ghost, int fields, float* out) {
    if (!out) {
        set_error("swe_hip_read_layer: null output");
        return SWE_HIP_EINVAL;
    }
    return layer_host_transfer(k, side, ghost, fields, out, nullptr);
}

int swe_hip_write_layer(swe_hip_block* k, int side, int ghost, int fields, const float* in) {
    if (!in) {
        set_error("swe_hip_write_layer: null input");
        return SWE_HIP_EINVAL;
    }
    return layer_host_transfer(k, side, ghost, fields, nullptr, in);
}

int swe_hip_solve_edges(int solver, int math, int64_t n, const float* hL, const float* hR, const float* huL,
                        const float* huR, const float* bL, const float* bR, float* hNetL, float* hNetR,
                        float* huNetL, float* huNetR, float* maxSpeed) {
    if (n < 0 || n > ((int64_t)1 << 28) || solver < 0 || solver > 3 || (math != 0 && math != 1)) {
        set_error("swe_hip_solve_edges: invalid n=%lld or solver=%d", (long long)n, solver);
        return SWE_HIP_EINVAL;
    }
    if (n == 0) return SWE_HIP_OK;
    const float* ins[6] = {hL, hR, huL, huR, bL, bR};
    float* outs[5] = {hNetL, hNetR, huNetL, huNetR, maxSpeed};
    for (auto p : ins)
        if (!p) {
            set_error("swe_hip_solve_edges: null input");
            return SWE_HIP_EINVAL;
        }
    for (auto p : outs)
        if (!p) {
            set_error("swe_hip_solve_edges: null output");
            return SWE_HIP_EINVAL;
        }
    float* din = nullptr;
    float* dout = nullptr;
    const size_t nb = sizeof(float) * (size_t)n;
    HIP_TRY(hipMalloc(&din, 6 * nb));
    hipError_t e = hipMalloc(&dout, 5 * nb);
    if (e != hipSuccess) {
        (void)hipFree(din);
        return hip_fail(e, "hipMalloc");
    }
    int rc = 0;
    for (int i = 0; i < 6 && e == hipSuccess; ++i) e = hipMemcpy(din + i * n, ins[i], nb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = launch_solve_edges(solver, math, n, din, dout, nullptr);
    }
    if (e == hipSuccess && rc == 0) e = hipDeviceSynchronize();
    for (int i = 0; i < 5 && e == hipSuccess && rc == 0; ++i)
        e = hipMemcpy(outs[i], dout + i * n, nb, hipMemcpyDeviceToHost);
    (void)hipFree(din);
    (void)hipFree(dout);
    if (e != hipSuccess) return hip_fail(e, "swe_hip_solve_edges");
    return rc;
}

int swe_hip_set_tuning(swe_hip_block* k, int strategy, int xseg) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (strategy != -1 && strategy != 0 && strategy != 1 && strategy != 2) {
        set_error("swe_hip_set_tuning: strategy must be 0, 1 or 2");
        return SWE_HIP_EINVAL;
    }
    if (xseg != -1 && xseg != 4 && xseg != 8 && xseg != 10 && xseg != 12 && xseg != 16 && xseg != 32) {
        set_error("swe_hip_set_tuning: xseg must be 4, 8, 10, 12, 16 or 32");
        return SWE_HIP_EINVAL;
    }
    if (strategy != -1) k->strategy = strategy;
    if (xseg != -1) {
        k->xseg = xseg;
        k->xsegSet = true;
    }
    return SWE_HIP_OK;
}

int swe_hip_set_quiet_skip(swe_hip_block* k, int enable) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (enable != 0 && enable != 1) {
        set_error("swe_hip_set_quiet_skip: enable must be 0 or 1");
        return SWE_HIP_EINVAL;
    }
    k->skipQuiet = enable != 0;
    return SWE_HIP_OK;
}

int swe_hip_set_lookahead(swe_hip_block* k, int enable) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = settle(k)) return r;
    if (enable != 0 && enable != 1) {
        set_error("swe_hip_set_lookahead: enable must be 0 or 1");
        return SWE_HIP_EINVAL;
    }
    k->lookahead = enable != 0;
    return SWE_HIP_OK;
}

int swe_hip_set_math(swe_hip_block* k, int math) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (math != 0 && math != 1) {
        set_error("swe_hip_set_math: math must be 0 (IEEE) or 1 (fast)");
        return SWE_HIP_EINVAL;
    }
    if (math == 1 && k->flavour != FLAVOUR_U) {
        set_error("swe_hip_set_math: the fast variant exists for flavour U only");
        return SWE_HIP_EINVAL;
    }
    if (math == 1 && k->mlp) {
        set_error("swe_hip_set_math: the learned middle state needs the IEEE variant");
        return SWE_HIP_EINVAL;
    }
    k->math = math;
    return SWE_HIP_OK;
}

int swe_hip_autotune(swe_hip_block* k, int steps, float* ms_per_step) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (steps < 2 || steps > 10000 || k->flavour != FLAVOUR_U) {
        set_error("swe_hip_autotune: needs 2 <= steps <= 10000 and flavour U");
        return SWE_HIP_EINVAL;
    }
    // snapshot of what a run changes and a later run reads: the current state
    // and the control block (the other buffer and the sums are scratch)
    const size_t bytes = field_elems(k) * sizeof(float);
    const int cur0 = k->cur;
    float* save[3] = {nullptr, nullptr, nullptr};
    DeviceCtl* saveCtl = nullptr;
    hipError_t e = hipSuccess;
    for (float*& p : save)
        if (e == hipSuccess) e = hipMalloc(&p, bytes);
    if (e == hipSuccess) e = hipMalloc(&saveCtl, sizeof(DeviceCtl));
    auto copy_state = [&](bool toSave) {
        k->cur = cur0;
        float* live[3] = {k->h[cur0], k->hu[cur0], k->hv[cur0]};
        for (int i = 0; i < 3 && e == hipSuccess; ++i)
            e = hipMemcpyAsync(toSave ? save[i] : live[i], toSave ? live[i] : save[i], bytes, hipMemcpyDeviceToDevice,
                               k->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(toSave ? saveCtl : k->ctl, toSave ? k->ctl : saveCtl, sizeof(DeviceCtl),
                               hipMemcpyDeviceToDevice, k->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(k->stream);
    };
    bool saved = false;
    if (e == hipSuccess) {
        copy_state(true);
        saved = e == hipSuccess;
    }
    const int widths[SWE_HIP_TUNE_WIDTHS] = {4, 8, 10, 12, 16, 32};
    const int oldXseg = k->xseg;
    int best = oldXseg, rc = 0;
    float bestMs = 0.f;
    // one untimed run first (first launches of the process, allocations)
    if (e == hipSuccess) rc = swe_hip_run_timed(k, 2, nullptr, nullptr, nullptr, nullptr);
    if (e == hipSuccess && rc == 0) copy_state(false);
    for (int i = 0; i < SWE_HIP_TUNE_WIDTHS && e == hipSuccess && rc == 0; ++i) {
        k->xseg = widths[i];
        float total = 0.f;
        rc = swe_hip_run_timed(k, steps, nullptr, &total, nullptr, nullptr);
        if (rc) break;
        const float per = total / steps;
        if (ms_per_step) ms_per_step[i] = per;
        if (i == 0 || per < bestMs) {
            bestMs = per;
            best = widths[i];
        }
        copy_state(false);
    }
    // the light solvers on HBM-resident blocks: the recompute strategy moves
    // 44 B per cell against the fused step's ~70 and can win at wide strips
    const int oldStrategy = k->strategy;
    int bestStrategy = oldStrategy;
    const int es = eff_solver(k);
    if (oldStrategy == 2 && (es == SOLVER_HLLE || es == SOLVER_FWAVE)) {
        k->strategy = 0;
        for (int xs : {16, 32}) {
            if (e != hipSuccess || rc != 0) break;
            k->xseg = xs;
            float total = 0.f;
            rc = swe_hip_run_timed(k, steps, nullptr, &total, nullptr, nullptr);
            if (rc) break;
            if (total / steps < bestMs) {
                bestMs = total / steps;
                best = xs;
                bestStrategy = 0;
            }
            copy_state(false);
        }
        k->strategy = oldStrategy;
    }
    // a failed timed run leaves the block half-advanced: put the snapshot back
    if (saved && (rc || e != hipSuccess)) {
        const hipError_t e0 = e;
        e = hipSuccess;
        copy_state(false);
        if (e0 != hipSuccess) e = e0;
    }
    for (float* p : save)
        if (p) (void)hipFree(p);
    if (saveCtl) (void)hipFree(saveCtl);
    k->cur = cur0;
    if (rc || e != hipSuccess) {
        k->xseg = oldXseg;
        k->strategy = oldStrategy;
        return rc ? rc : hip_fail(e, "swe_hip_autotune");
    }
    k->xseg = best;
    k->strategy = bestStrategy;
    k->xsegSet = true;
    return SWE_HIP_OK;
}

int swe_hip_get_tuning(swe_hip_block* k, int* strategy, int* xseg) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (strategy) *strategy = k->strategy;
    if (xseg) *xseg = k->xseg;
    return SWE_HIP_OK;
}

int swe_hip_set_surrogate(swe_hip_block* k, swe_hip_mlp* m) {
    if (int r = check(k)) return r;
    std::lock_guard<std::recursive_mutex> lock(k->mu);
    if (int r = touch(k)) return r;
    if (m && (k->solver != SOLVER_AUGRIE || k->flavour != FLAVOUR_U)) {
        set_error("swe_hip_set_surrogate: the learned middle state serves flavour-U AugRie blocks only");
        return SWE_HIP_EINVAL;
    }
    if (m && k->math != 0) {
        set_error("swe_hip_set_surrogate: the learned middle state needs the IEEE variant");
        return SWE_HIP_EINVAL;
    }
    const unsigned* frags = m ? swe::mlp_frags(m) : nullptr;
    HIP_TRY(hipStreamSynchronize(k->stream));
    HIP_TRY(hipMemcpy(&k->ctl->mlp, &frags, sizeof(frags), hipMemcpyHostToDevice));
    if (k->net != m) {
        if (k->net) swe::mlp_detach(k->net, k->stream);
        if (m) swe::mlp_attach(m, k->stream);
        k->net = m;
    }
    k->mlp = frags;
    if (!k->xsegSet) k->xseg = default_xseg(k->nx, k->ny, frags ? SOLVER_AUGRIE_DEFER : k->solver);
    return SWE_HIP_OK;
}

}  // extern "C"

// ---- the entry point that allocates nothing ---------------------------------
// Scratch of swe_hip_run_noalloc: the control block, both sets of net-update
// sums, the quiet-tile record and the plan's lists.  One copy per GPU, held
// until the process ends; calls that overlap on one GPU race on it.
static __device__ __align__(256) unsigned char scratch[SWE_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 256;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of the current GPU's scratch array, looked up once per
// GPU with that GPU current.  In host code the name `scratch` stands for no
// device address: only this pointer may reach kernels and copies.
int noalloc_scratch(unsigned char** out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("swe_hip_run_noalloc: device %d out of range", dev);
        return SWE_HIP_EDEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Scratch bytes a block of nx x ny cells needs, and the offsets of its parts:
// control block, sums A, sums B (4 fields each), quiet flags, tile speeds,
// plan list, plan counts.
size_t noalloc_layout(const swe_hip_block& k, size_t off[7]) {
    const size_t field = align256(field_elems(&k) * sizeof(float));
    const size_t tiles = max_tiles(&k);
    const size_t groups = plan_groups_max(&k);
    size_t at = 0;
    off[0] = at;
    at += align256(sizeof(DeviceCtl));
    off[1] = at;
    at += 4 * field;
    off[2] = at;
    at += 4 * field;
    off[3] = at;
    at += align256(2 * tiles * sizeof(unsigned));
    off[4] = at;
    at += align256(2 * tiles * sizeof(float));
    off[5] = at;
    at += align256(groups * kPlanGroup * kPlanPieces * sizeof(unsigned));
    off[6] = at;
    at += align256(groups * sizeof(unsigned));
    return at;
}

__global__ void copy_state_kernel(const float* __restrict__ h, const float* __restrict__ hu,
                                  const float* __restrict__ hv, float* __restrict__ oh, float* __restrict__ ohu,
                                  float* __restrict__ ohv, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        oh[i] = h[i];
        ohu[i] = hu[i];
        ohv[i] = hv[i];
    }
}

}  // namespace

extern "C" {

int swe_hip_run_noalloc(int nx, int ny, float dx, float dy, int solver, const int* boundary, float* h, float* hu,
                        float* hv, const float* b, float* h2, float* hu2, float* hv2, int nsteps, float* dt_log,
                        void* stream) {
    if (nx < 1 || ny < 1 || !(dx > 0.f) || !(dy > 0.f) || solver < 0 || solver > 3 || !boundary || !h || !hu ||
        !hv || !b || !h2 || !hu2 || !hv2 || nsteps < 0) {
        set_error("swe_hip_run_noalloc: invalid argument (nx=%d ny=%d dx=%g dy=%g solver=%d nsteps=%d)", nx, ny, dx,
                  dy, solver, nsteps);
        return SWE_HIP_EINVAL;
    }
    for (int i = 0; i < 4; ++i)
        if (boundary[i] < OUTFLOW || boundary[i] > PASSIVE || boundary[i] == INFLOW) {
            set_error("swe_hip_run_noalloc: boundary type %d of side %d is not supported", boundary[i], i);
            return SWE_HIP_EINVAL;
        }
    if ((size_t)(nx + 2) * (size_t)(ny + 2) > ((size_t)1 << 31)) {
        set_error("swe_hip_run_noalloc: block too large");
        return SWE_HIP_EINVAL;
    }
    if (nsteps == 0) return SWE_HIP_OK;
    // a block description on the stack whose device parts all point into the
    // caller's arrays and the scratch array, so the loop allocates nothing
    swe_hip_block k;
    k.nx = nx;
    k.ny = ny;
    k.pitch = ny + 2;
    k.dx = dx;
    k.dy = dy;
    k.solver = solver;
    k.flavour = FLAVOUR_U;
    k.strategy = 2;
    size_t off[7];
    const size_t need = noalloc_layout(k, off);
    if (need > (size_t)SWE_HIP_NOALLOC_SCRATCH_BYTES) {
        set_error("swe_hip_run_noalloc: a %d x %d block needs %zu bytes of scratch, more than the %zu this entry "
                  "point holds (SWE_HIP_NOALLOC_SCRATCH_BYTES); use swe_hip_run",
                  nx, ny, need, (size_t)SWE_HIP_NOALLOC_SCRATCH_BYTES);
        return SWE_HIP_EINVAL;
    }
    k.xseg = default_xseg(nx, ny, solver);
    unsigned char* base = nullptr;
    if (int r = noalloc_scratch(&base)) return r;
    const size_t field = align256(field_elems(&k) * sizeof(float));
    for (int i = 0; i < 4; ++i) {
        k.sums[i] = reinterpret_cast<float*>(base + off[1] + i * field);
        k.sums2[i] = reinterpret_cast<float*>(base + off[2] + i * field);
    }
    k.ctl = reinterpret_cast<DeviceCtl*>(base + off[0]);
    k.quiet = reinterpret_cast<unsigned*>(base + off[3]);
    k.tspeed = reinterpret_cast<float*>(base + off[4]);
    k.tlist = reinterpret_cast<unsigned*>(base + off[5]);
    k.tcount = reinterpret_cast<unsigned*>(base + off[6]);
    k.h[0] = h;
    k.hu[0] = hu;
    k.hv[0] = hv;
    k.h[1] = h2;
    k.hu[1] = hu2;
    k.hv[1] = hv2;
    k.b = const_cast<float*>(b);
    k.cur = 0;
    k.dtlog = dt_log;
    k.dtlogCap = 1 << 30;
    k.stream = static_cast<hipStream_t>(stream);
    for (int i = 0; i < 4; ++i) k.bi.type[i] = boundary[i];
    int rc = run_u_device_loop(&k, nsteps, nullptr, nullptr, nullptr, false);
    if (rc == 0 && k.cur != 0) {  // the state ended in the second buffer: bring it home
        const size_t n = field_elems(&k);
        hipLaunchKernelGGL(copy_state_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0,
                           k.stream, h2, hu2, hv2, h, hu, hv, n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_fail(e, "copy_state_kernel launch");
    }
    // k is a view: nothing of it is freed (its destructor-free struct owns nothing)
    return rc;
}

}  // extern "C"
