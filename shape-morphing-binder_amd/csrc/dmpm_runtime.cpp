// Host runtime behind the C ABI (include/dmpm.h).
//
// A dmpm_sim is the device-resident counterpart of the reference's CompGraph
// (CompGraph.h:20-99): a sequence of layers (particle states), the grid
// scratch of one step, the adjoint carry, per-layer controls and Adam state.
// Unlike the reference's store-every-step tape (one full PointCloud and Grid
// per layer, CompGraph.cpp:30-40), only every K-th layer (and the last) is
// stored; the others are recomputed segment by segment during the backward
// pass.  The grid is never stored: each backward step re-runs that step's P2G.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <map>
#include <vector>

#include "../../include/dmpm.h"
#include "dmpm_kernels.h"

using namespace dmpm;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw Error(DMPM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
    } while (0)

template <typename F>
int guard(F&& f) {
    try {
        f();
        return DMPM_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return DMPM_ERR_ARG;
    }
}

// Process-wide Adam step counter of OptimizeSingleTimestep, which the
// reference keeps in a function-local static (CompGraph.cpp:396).
int g_single_timestep_adam_t = 0;

long long round_up(long long n, long long a) { return (n + a - 1) / a * a; }

}  // namespace

void dmpm::set_last_error(const std::string& msg) { g_err = msg; }

struct dmpm_sim {
    dmpm_config cfg{};
    Phys ph{};
    Prec pr = PREC_F32;
    size_t esz = 4;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    int n = 0;
    long long ld = 0;
    long long G = 0, gld = 0;
    int L = 1;
    int K = 1;

    // device memory bookkeeping: allocations in use, and freed ones kept for
    // reuse (size -> blocks).  hipMalloc / hipFree of the hundreds of layer
    // buffers a set-up allocates cost milliseconds each (hipFree also waits
    // for the device); a sim that is set up again with the same particle
    // count -- every call of a differentiable rollout -- takes them from the
    // pool instead.  mem_cur counts both (device memory the sim holds).
    size_t mem_cur = 0, mem_peak = 0, pooled = 0;
    std::vector<std::pair<void*, size_t>> allocs;
    std::multimap<size_t, void*> pool;

    int* perm = nullptr;   // internal index -> caller index
    int* bst = nullptr;    // particle block table of the window kernels (nblk + 1 starts)
    int nblk = 0;
    void* stat = nullptr;  // m, vol, lam, mu

    // ---- storage orders of episode segments (dmpm_sort_segments) ----
    // The layers [ob[s], ob[s+1]) of segment s keep every per-particle array
    // (state, control, Adam state, control gradient) in storage order s, each
    // segment sorted by its own middle layer; the adjoint carry, the last
    // layer's seed and render gradients follow the order of the layer they
    // belong to.  Order 0 is perm / bst / nblk / stat above; orders 1.. are in
    // oseg.  Without segment orders (the default) ob is empty.
    struct OrderSeg {
        int* perm = nullptr;
        int* bst = nullptr;
        int nblk = 0;
        void* stat = nullptr;
        int* fwdq = nullptr;    // slot j of this order <- index in the previous order
        int* bwdq = nullptr;    // slot j of the previous order <- index in this order
        int* zq = nullptr;      // slot j of order 0 <- index in this order
    };
    std::vector<OrderSeg> oseg;
    std::vector<int> ob;
    void* ctl_x = nullptr;  // a segment's first control staged in the previous order
    int nord() const { return 1 + (int)oseg.size(); }
    int ord_of(int t) const {
        int o = 0;
        while (o + 1 < nord() && t >= ob[o + 1]) ++o;
        return o;
    }
    const int* perm_of(int o) const { return o == 0 ? perm : oseg[o - 1].perm; }
    void* stat_of(int o) const { return o == 0 ? stat : oseg[o - 1].stat; }
    std::vector<void*> layer;  // stored layer states (nullptr = not stored)
    std::vector<char> valid;   // layer computed (stored ones only)
    // pend[t]: stored layer t holds a copy of layer 0 that set-up has not
    // made yet (it is made when the layer is first read; the forward pass
    // overwrites it first in the usual flow).  Only in one storage order.
    std::vector<char> pend;
    void materialize(int t) {
        if (t < (int)pend.size() && pend[t]) {
            copy_elems(pr, layer[t], layer[0], NSTATE * ld, stream);
            pend[t] = 0;
        }
    }
    void materialize_all() {
        for (int t = 1; t < (int)pend.size(); ++t) materialize(t);
    }
    std::vector<void*> ctl;    // dFc per layer (nullptr = zero)
    std::vector<void*> am, av, avmax;  // Adam moments per layer (PointCloud momentum/vector/vector_max)
    std::vector<char> want_gctl;
    std::vector<void*> gctl;   // dL/d dFc per layer
    std::vector<void*> seg;    // K segment buffers (recomputed states)
    std::vector<void*> seg_grid;  // grids (m, p) of the segment's layers
    int seg_base = -1;
    // With K = 1 (and recompute_grids off) the forward pass keeps every
    // layer's P2G grid (m, p) for the backward pass, as the reference's
    // CompGraph keeps each layer's Grid: lgrid[t] is layer t's grid,
    // gvalid[t] whether it matches layer t's state and control.
    std::vector<void*> lgrid;
    std::vector<char> gvalid;
    // fresh[t]: layer t's state (stored, or in the current segment) is the
    // forward step of layer t-1 under the current control of layer t-1 and
    // the current parameters.  The adjoint of step t-1 then reads v' and C'
    // from it instead of recomputing them by a G2P.
    std::vector<char> fresh;
    bool keep_grids() const { return K <= 1 && !cfg.recompute_grids; }
    // A control (or a parameter the P2G reads) changed: layer t's grid (all
    // grids for t < 0), any recomputed segment and the freshness of layer
    // t+1 (of every layer for t < 0) no longer hold.
    void grids_stale(int t) {
        seg_base = -1;
        if (t < 0) {
            std::fill(gvalid.begin(), gvalid.end(), 0);
            std::fill(fresh.begin(), fresh.end(), 0);
        } else {
            if (t < (int)gvalid.size()) gvalid[t] = 0;
            if (t + 1 < (int)fresh.size()) fresh[t + 1] = 0;
        }
    }
    void* scratch[2] = {nullptr, nullptr};
    void* grid = nullptr;  // m, p0, p1, p2 per node (gld nodes)
    void* grid2 = nullptr; // second grid for the fused forward step
    unsigned long long* wstat = nullptr;  // scatter window counters (device, 2 entries)
    void* det = nullptr;   // deterministic-mode scratch
    void* det_buf() {
        if (!det) {
            const size_t b = det_scratch_bytes(n, cfg.dims, (int)esz);
            det = dalloc(b);
        }
        return det;
    }
    void* ggv = nullptr;   // dL/dv_i per node (4 values, the last unused)
    bool count_windows = false;  // window counters (wstat) updated by the scatter launches
    void* ggv2 = nullptr;  // as ggv: dL/dv of the previous step, written by the fused adjoint
    void* ggv3 = nullptr;  // as ggv: the dL/dv grid a fused adjoint launch clears for the next one
    void* carry[2] = {nullptr, nullptr};
    int carry_cur = 0;
    int carry_layer = -1;   // layer whose adjoint the current carry holds
    void* adj_last = nullptr;  // adjoint of the last layer (loss / render / seed)
    void* gctl_stop = nullptr; // dL/d dFc of the control layer of the last backward
    void* target_m = nullptr;
    void* mgrid = nullptr;  // node masses of the end-layer loss (one per node)
    void* dldm = nullptr;
    double* partial = nullptr;  // reduction partials
    double* red = nullptr;      // reduction outputs (8 doubles)
    int red_blocks = 1024;
    // render gradients, internal order, planes: F (9) + x (3)
    void* rg = nullptr;
    bool has_rg = false;
    int rg_n = 0;
    double rg_normF = 0, rg_normx = 0;
    double* rgn = nullptr;        // device: squared norms of the last set (F, x)
    bool rg_norms_host = true;    // rg_normF/x hold them
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double fwd_ms = 0, bwd_ms = 0;
    // per-kernel event timing (dmpm_kernel_timer): pairs of events per launch
    static constexpr int NKIND = 7;  // p2g, g2p, adj_scatter, adj_gather, grid clear, fused step, fused adjoint
    bool ktimer = false;
    bool capturing = false;  // inside dmpm_graph_capture: no timing events
    // captured graphs and batched episodes alive that hold this sim's device
    // pointers (shared with them, so that one destroyed after the sim still
    // counts down safely)
    std::shared_ptr<int> captures = std::make_shared<int>(0);
    // cleared when the sim is destroyed; a graph or batch refuses to launch
    // once any sim it captured is gone
    std::shared_ptr<bool> alive = std::make_shared<bool>(true);
    std::vector<hipEvent_t> kpool;
    size_t knext = 0;
    std::vector<std::pair<int, size_t>> krec;  // (kind, index of start event)
    int koverflow = 0;

    // A captured graph or batched episode holds this sim's device pointers
    // (layers, grids, carries, controls, render gradients, Adam-free
    // buffers) and the pointers and sizes baked into its launches.  Calls
    // that would free, move or resize such a buffer, or change a parameter
    // baked into a launch, are refused while one is alive.
    bool captured() const { return *captures > 0; }
    void refuse_if_captured(const char* what) const {
        if (captured())
            throw Error(DMPM_ERR_STATE, std::string(what) + " is refused while a captured graph or batched episode "
                                                            "holds this sim's buffers; destroy it first");
    }

    // Host bookkeeping after an episode (which layers, grids and segment
    // buffers hold results, where the adjoint carry is).  A capture runs the
    // bookkeeping of the episode it records; each launch of the capture
    // restores it, since a replay leaves the device in that same state
    // whatever eager calls ran in between.
    struct Book {
        std::vector<char> valid, fresh, gvalid;
        int seg_base = -1, carry_cur = 0, carry_layer = -1;
        void* seg0 = nullptr;
    };
    Book book() const {
        Book b;
        b.valid = valid;
        b.fresh = fresh;
        b.gvalid = gvalid;
        b.seg_base = seg_base;
        b.carry_cur = carry_cur;
        b.carry_layer = carry_layer;
        b.seg0 = seg.empty() ? nullptr : seg[0];
        return b;
    }
    void restore(const Book& b) {
        valid = b.valid;
        fresh = b.fresh;
        gvalid = b.gvalid;
        seg_base = b.seg_base;
        carry_cur = b.carry_cur;
        carry_layer = b.carry_layer;
        if (!seg.empty()) seg[0] = b.seg0;
    }

    ~dmpm_sim() {
        *alive = false;
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& a : allocs) (void)hipFree(a.first);
        for (auto& a : pool) (void)hipFree(a.second);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : kpool) (void)hipEventDestroy(e);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
    }

    // ---- memory ----
    // A freed block is only handed out again for work queued on the sim's
    // stream after the free, which runs after every earlier use of it.
    void* dalloc(size_t bytes) {
        if (bytes == 0) bytes = 16;
        void* p = nullptr;
        auto it = pool.find(bytes);
        if (it != pool.end()) {
            p = it->second;
            pool.erase(it);
            pooled -= bytes;
        } else {
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess && !pool.empty()) {
                (void)hipGetLastError();
                release_pool();
                e = hipMalloc(&p, bytes);
            }
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw Error(DMPM_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " +
                                                hipGetErrorString(e));
            }
            mem_cur += bytes;
            mem_peak = std::max(mem_peak, mem_cur);
        }
        allocs.emplace_back(p, bytes);
        return p;
    }
    void dfree(void*& p) {
        if (!p) return;
        for (size_t i = allocs.size(); i-- > 0;)
            if (allocs[i].first == p) {
                pool.emplace(allocs[i].second, p);
                pooled += allocs[i].second;
                allocs.erase(allocs.begin() + i);
                break;
            }
        p = nullptr;
    }
    // Give the pooled blocks back to the device.
    void release_pool() {
        if (pool.empty()) return;
        for (auto& a : pool) (void)hipFree(a.second);
        mem_cur -= pooled;
        pooled = 0;
        pool.clear();
    }
    size_t state_bytes() const { return (size_t)NSTATE * ld * esz; }
    size_t ctl_bytes() const { return (size_t)9 * ld * esz; }
    void zero(void* p, long long elems) { fill_zero(pr, p, elems, stream); }

    // ---- launch descriptor ----
    // of the window kernels over the particles of a layer stored in order o
    Launch launch(int o) const {
        Launch Lc = launch();
        if (o > 0) {
            Lc.bst = oseg[o - 1].bst;
            Lc.nblk = oseg[o - 1].nblk;
        }
        return Lc;
    }
    Launch launch() const {
        Launch Lc{};
        Lc.ph = ph;
        Lc.n = n;
        Lc.ld = ld;
        Lc.gld = gld;
        Lc.wstat = count_windows ? wstat : nullptr;
        Lc.bst = bst;
        Lc.nblk = nblk;
        // default LDS budget: the windows of 256-particle blocks cut at 704
        // stencil nodes fit 30 KiB with FP32 state.  The fused step holds 94
        // VGPRs (five waves per SIMD), so five blocks should share a CU's
        // 160 KiB; with the kernels' few hundred bytes of static LDS, a
        // 32 KiB (or 31 KiB) budget leaves room for four only (measured on
        // the 1M-particle workload: 112 µs per step at 32 KiB, 105 µs at
        // 30 KiB).  FP64 staging needs 48 KiB.
        Lc.lds_bytes = cfg.lds_bytes > 0 ? cfg.lds_bytes : (pr == PREC_F32 ? 30 : 48) * 1024;
        Lc.block = cfg.block > 0 ? cfg.block : 256;
        return Lc;
    }

    // Run `launch` between two pooled events when the kernel timer is on.
    template <typename Fn>
    void timed(int kind, Fn&& launch) {
        if (!ktimer) {
            launch();
            return;
        }
        if (knext + 2 > kpool.size()) {
            ++koverflow;
            launch();
            return;
        }
        const size_t i = knext;
        knext += 2;
        HIPCHK(hipEventRecord(kpool[i], stream));
        launch();
        HIPCHK(hipEventRecord(kpool[i + 1], stream));
        krec.emplace_back(kind, i);
    }

    // ---- GPU operations of the forward and backward passes ----
    // Each either runs now or, while an episode is being recorded for a
    // batched launch (rec != null), is appended to rec instead.
    struct Op {
        int kind;
        Launch L;
        void* dst;
        const void* src;
        long long n;
        Launch L2;  // second launch of a fused adjoint op
    };
    std::vector<Op>* rec = nullptr;

    void op_zero(void* p, long long elems) {
        if (rec) {
            rec->push_back(Op{OP_ZERO, Launch{}, p, nullptr, elems, Launch{}});
            return;
        }
        timed(4, [&] { zero(p, elems); });
    }
    void op_copy(void* dst, const void* src, long long elems) {
        if (rec) {
            rec->push_back(Op{OP_COPY, Launch{}, dst, src, elems, Launch{}});
            return;
        }
        copy_elems(pr, dst, src, elems, stream);
    }
    // dst += src over `elems` elements of the sim's precision
    void op_accum(void* dst, const void* src, long long elems) {
        if (rec) {
            rec->push_back(Op{OP_ACCUM, Launch{}, dst, src, elems, Launch{}});
            return;
        }
        accumulate_carry(pr, src, pr == PREC_F64, dst, 0, 1, nullptr, (int)elems, elems, stream);
    }
    void op_kernel(int kind, const Launch& Lc) {
        if (rec) {
            rec->push_back(Op{kind, Lc, nullptr, nullptr, 0, Launch{}});
            return;
        }
        switch (kind) {
            case OP_P2G: timed(0, [&] { p2g(pr, Lc, false, stream); }); break;
            case OP_P2G_MASS: timed(0, [&] { p2g(pr, Lc, true, stream); }); break;
            case OP_G2P: timed(1, [&] { g2p(pr, Lc, stream); }); break;
            case OP_STEP: timed(5, [&] { dmpm::step(pr, Lc, stream); }); break;
            case OP_ADJ_SCATTER: timed(2, [&] { adj_scatter(pr, Lc, stream); }); break;
            case OP_ADJ_GATHER: timed(3, [&] { adj_gather(pr, Lc, stream); }); break;
            default: throw Error(DMPM_ERR_ARG, "unknown kernel op");
        }
    }
    // The adjoint gather of step t (Lc) fused with the adjoint scatter of
    // step t-1 (Lp).
    void op_adj_fused(const Launch& Lc, const Launch& Lp) {
        if (rec) {
            rec->push_back(Op{OP_ADJ_FUSED, Lc, nullptr, nullptr, 0, Lp});
            return;
        }
        timed(6, [&] { adj_fused(pr, Lc, Lp, stream); });
    }

    bool stored(int t) const { return K <= 1 || t % K == 0 || t == L - 1; }

    // (Re)allocate layer storage for L layers.
    void alloc_layers() {
        for (int t = 1; t < (int)layer.size(); ++t) dfree(layer[t]);
        // seg[0] aliases a stored layer; only seg[1..] are owned
        for (size_t j = 1; j < seg.size(); ++j) dfree(seg[j]);
        for (auto& g : seg_grid) dfree(g);
        for (auto& g : lgrid) dfree(g);
        seg.clear();
        seg_grid.clear();
        lgrid.clear();
        gvalid.clear();
        seg_base = -1;
        layer.resize(L, nullptr);
        valid.assign(L, 0);
        valid[0] = 1;
        pend.assign(L, 0);
        fresh.assign(L, 0);
        for (int t = 1; t < L; ++t)
            if (stored(t)) layer[t] = dalloc(state_bytes());
        if (K > 1 && L > 1) {
            const int k = std::min(K, L);
            seg.resize(k, nullptr);
            seg_grid.resize(k, nullptr);
            for (int j = 1; j < k; ++j) seg[j] = dalloc(state_bytes());
            for (int j = 0; j < k; ++j) seg_grid[j] = dalloc(4 * gld * esz);
        }
        if (keep_grids() && L > 1) {
            lgrid.resize(L - 1, nullptr);
            gvalid.assign(L - 1, 0);
            for (auto& g : lgrid) g = dalloc(4 * gld * esz);
        }
        for (auto& p : scratch)
            if (!p) p = dalloc(state_bytes());
    }

    // Recompute the K-layer segment holding layer t (K > 1): the states of
    // layers c+1..top from the stored checkpoint c, and the grids (m, p) of
    // layers c..top, which the backward steps of the segment read.
    void ensure_segment(int t) {
        const int c = (t / K) * K;
        if (seg_base == c) return;
        if (!valid[c]) throw Error(DMPM_ERR_STATE, "checkpoint layer " + std::to_string(c) + " missing");
        materialize(c);
        const int top = std::min(c + K - 1, L - 2);
        seg_base = -1;
        seg[0] = layer[c];
        p2g_into(seg[0], c, seg_grid[0]);
        // step j also clears the grid step j+1 accumulates into
        bool g1_clear = false;
        for (int j = c; j < top; ++j) {
            void* ahead = j + 2 <= top ? seg_grid[j - c + 2] : nullptr;
            g1_clear = fused_step(seg[j - c], j, seg_grid[j - c], seg[j - c + 1], seg_grid[j - c + 1], g1_clear, ahead);
            fresh[j + 1] = 1;
        }
        seg_base = c;
    }

    // Device pointer of the state of layer t, recomputing it from the nearest
    // stored layer below when it is not stored.
    const void* state(int t) {
        if (stored(t)) {
            if (!valid[t]) throw Error(DMPM_ERR_STATE, "layer " + std::to_string(t) + " has not been computed");
            materialize(t);
            return layer[t];
        }
        ensure_segment(t);
        return seg[t - seg_base];
    }

    // State of layer t if it is in device memory now (stored and computed,
    // or in the current segment) without a recompute, else null.
    const void* resident_state(int t) const {
        if (t < 0 || t >= L) return nullptr;
        if (K > 1) {
            if (seg_base < 0 || t < seg_base || t >= seg_base + K || t > L - 2) return nullptr;
            return t == seg_base ? layer[t] : seg[t - seg_base];
        }
        return valid[t] && !pend[t] ? layer[t] : nullptr;
    }

    // Zero `g` and P2G layer t's state `in` into it.
    void p2g_into(const void* in, int t, void* g) {
        Launch Lc = launch(ord_of(t));
        Lc.st = in;
        Lc.stat = stat_of(ord_of(t));
        Lc.ctl = ctl[t];
        Lc.grid = g;
        if (cfg.deterministic) {
            if (rec) throw Error(DMPM_ERR_STATE, "deterministic mode cannot be batched");
            timed(0, [&] { det_p2g(pr, Lc, false, det_buf(), stream); });
            return;
        }
        op_zero(g, 4 * gld);
        op_kernel(OP_P2G, Lc);
    }

    // G2P of layer t (`in`, grid `g`) -> `out`; no P2G of the result.
    void g2p_only(const void* in, int t, void* g, void* out) {
        Launch Lc = launch(ord_of(t));
        Lc.st = in;
        Lc.stat = stat_of(ord_of(t));
        Lc.ctl = ctl[t];
        Lc.grid = g;
        Lc.out = out;
        op_kernel(OP_G2P, Lc);
    }

    // Layer t (`in`, grid `g`) -> layer t+1 (`out`) and its grid `g1`.
    // g1_clear: g1 was cleared by the previous launch.  ahead: the grid of
    // the next step, cleared by this launch's blocks when the fused kernel
    // runs (returns whether it did).
    // `ctl1`, when given, is layer t+1's control in layer t's storage order,
    // and `opos` the row of each particle in layer t+1's order (a step into
    // a segment of another order).
    bool fused_step(const void* in, int t, void* g, void* out, void* g1, bool g1_clear = false,
                    void* ahead = nullptr, const void* const* ctl1 = nullptr, const int* opos = nullptr) {
        Launch Lc = launch(ord_of(t));
        Lc.st = in;
        Lc.stat = stat_of(ord_of(t));
        Lc.ctl = ctl[t];
        Lc.grid = g;
        Lc.out = out;
        Lc.ctl1 = ctl1 ? *ctl1 : t + 1 < L ? ctl[t + 1] : nullptr;
        Lc.opos = opos;
        Lc.grid1 = g1;
        if (cfg.unfused || cfg.deterministic) {
            op_kernel(OP_G2P, Lc);
            p2g_into(out, t + 1, g1);
            return false;
        }
        if (!g1_clear) op_zero(g1, 4 * gld);
        if (ahead) {
            Lc.zero = ahead;
            Lc.zero_bytes = 4 * gld * (long long)esz;
        }
        op_kernel(OP_STEP, Lc);
        return ahead != nullptr;
    }

    void forward(int start, bool zero_controls = true) {
        if (start < 0 || start >= L) throw Error(DMPM_ERR_INDEX, "start layer out of range");
        const void* cur = state(start);
        // The reference zeroes dFc of every layer it writes (ForwardSimulation.cpp:156);
        // a rollout with per-step controls keeps them.  While a capture holds
        // the control buffers they are zeroed in place instead of freed.
        if (zero_controls)
            for (int t = start + 1; t < L; ++t) {
                if (!ctl[t]) continue;
                if (captured()) {
                    zero(ctl[t], 9 * ld);
                    grids_stale(t);
                } else {
                    dfree(ctl[t]);
                }
            }
        // The segment the backward pass enters first (holding layer L-2) is
        // written straight into the segment buffers -- states and grids --
        // when this pass computes all of it, so the backward pass need not
        // recompute it.  `cur` may live in a segment buffer of an earlier
        // segment; it is read by the first step only, before any layer of the
        // last segment is written.
        const int c_last = K > 1 && L >= 2 ? ((L - 2) / K) * K : -1;
        const bool keep = c_last >= 0 && start <= c_last;
        const bool kg = keep_grids();
        if (start + 1 < L) {
            if (!capturing && !rec) HIPCHK(hipEventRecord(ev[0], stream));
            void* ga = kg ? lgrid[start] : keep && start == c_last ? seg_grid[0] : grid;
            p2g_into(cur, start, ga);
            if (kg) gvalid[start] = 1;
            int sc = 0;
            // Where step t+1 accumulates into a grid of its own (every
            // layer's grid kept, or the segment buffers of the last segment)
            // step t's launch clears it (neither read nor written by step t),
            // and step t+1 needs no clear launch.  Not with the two grids
            // used in turn, one of which step t reads.
            auto target = [&](int u) -> void* {
                if (kg) return lgrid[u + 1];
                if (keep && u + 1 >= c_last) return seg_grid[u + 1 - c_last];
                return nullptr;
            };
            bool gb_clear = false;
            for (int t = start; t < L - 1; ++t) {
                const bool in_last = keep && t + 1 >= c_last;
                void* out;
                if (stored(t + 1)) {
                    out = layer[t + 1];
                } else if (in_last) {
                    out = seg[t + 1 - c_last];
                } else {
                    out = scratch[sc];
                    sc ^= 1;
                }
                const int o1 = ord_of(t + 1);
                if (o1 != ord_of(t) && t + 1 < L - 1 && !cfg.unfused) {
                    // Into a segment of another storage order: the fused step
                    // in the old order (G2P of layer t, P2G of layer t+1 with
                    // its control staged in the old order) writes layer t+1
                    // straight into its rows in the new order.
                    OrderSeg& os = oseg[o1 - 1];
                    const void* c1 = nullptr;
                    if (ctl[t + 1]) {
                        if (!ctl_x) ctl_x = dalloc(ctl_bytes());
                        permute_planes(pr, ctl[t + 1], ctl_x, os.bwdq, n, 9, ld, stream);
                        c1 = ctl_x;
                    }
                    void* gb = kg        ? lgrid[t + 1]
                               : in_last ? seg_grid[t + 1 - c_last]
                                         : (ga == grid ? grid2 : grid);
                    void* ahead = t + 2 < L - 1 ? target(t + 1) : nullptr;
                    gb_clear = fused_step(cur, t, ga, out, gb, gb_clear, ahead, &c1, os.bwdq);
                    if (kg) gvalid[t + 1] = 1;
                    ga = gb;
                } else if (o1 != ord_of(t)) {
                    // the same without the fused step: G2P in the old order
                    // into the scratch state, permuted into the new order,
                    // and the layer's P2G in the new order
                    OrderSeg& os = oseg[o1 - 1];
                    void* tmp = scratch[sc];
                    g2p_only(cur, t, ga, tmp);
                    permute_planes(pr, tmp, out, os.fwdq, n, NSTATE, ld, stream);
                    if (t + 1 < L - 1) {
                        void* gb = kg        ? lgrid[t + 1]
                                   : in_last ? seg_grid[t + 1 - c_last]
                                             : (ga == grid ? grid2 : grid);
                        p2g_into(out, t + 1, gb);
                        if (kg) gvalid[t + 1] = 1;
                        ga = gb;
                        gb_clear = false;
                    }
                } else if (t + 1 < L - 1) {
                    void* gb = kg        ? lgrid[t + 1]
                               : in_last ? seg_grid[t + 1 - c_last]
                                         : (ga == grid ? grid2 : grid);
                    void* ahead = t + 2 < L - 1 ? target(t + 1) : nullptr;
                    gb_clear = fused_step(cur, t, ga, out, gb, gb_clear, ahead);
                    if (kg) gvalid[t + 1] = 1;
                    ga = gb;
                } else {
                    g2p_only(cur, t, ga, out);
                }
                if (stored(t + 1)) {
                    valid[t + 1] = 1;
                    pend[t + 1] = 0;
                }
                fresh[t + 1] = 1;
                cur = out;
            }
            if (!capturing && !rec) HIPCHK(hipEventRecord(ev[1], stream));
        }
        seg_base = -1;
        if (keep) {
            seg[0] = layer[c_last];
            seg_base = c_last;
        }
    }

    void backward(int stop) {
        if (stop < 0 || stop > L - 1) throw Error(DMPM_ERR_INDEX, "control layer out of range");
        // render gradients are added to the last layer on every call (CompGraph.cpp:148-197)
        if (has_rg && rg_n == n) {
            accumulate_planes(rg, PF, 9);
            accumulate_planes((char*)rg + 9 * ld * esz, PX, 3);
        } else if (has_rg) {
            fprintf(stderr, "[WARN] Render gradient size mismatch: stored %d but layer has %d points. Skipping injection.\n",
                    rg_n, n);
        }
        op_copy(carry[carry_cur], adj_last, NSTATE * ld);
        carry_layer = L - 1;
        if (!capturing && !rec) HIPCHK(hipEventRecord(ev[2], stream));
        // dL/dv grid of the step being processed, and the other one; a fused
        // launch (gather of step t + scatter of step t-1) leaves step t-1's
        // dL/dv in the other grid, so the next step skips its scatter.  A
        // fused launch also clears the third grid (last read by the previous
        // step's gather), which the next fused launch scatters into.
        void* gcur = ggv;
        void* gnext = ggv2;
        void* gfree = ggv3;
        bool gnext_clear = false;
        bool scattered = false;
        for (int t = L - 2; t >= stop; --t) {
            const void* st;
            void* g;
            const int o = ord_of(t);
            // Out of a segment of another storage order: the carry of layer
            // t+1 into this order.  The adjoint step of layer t+1 wrote it
            // there itself; only the seed of the last layer is permuted.
            if (ord_of(t + 1) != o && t + 1 == L - 1) {
                permute_planes(pr, carry[carry_cur], carry[carry_cur ^ 1], oseg[o].bwdq, n, NSTATE, ld, stream);
                carry_cur ^= 1;
            }
            // layer t+1 when it is the forward step of layer t as it stands
            // now (recomputed segment layers always are)
            const void* nxt = nullptr;
            if (K > 1) {
                ensure_segment(t);
                st = t == seg_base ? layer[t] : seg[t - seg_base];
                g = seg_grid[t - seg_base];
                if (fresh[t + 1]) {
                    if (!stored(t + 1)) nxt = seg[t + 1 - seg_base];
                    else if (valid[t + 1]) nxt = layer[t + 1];
                }
            } else if (keep_grids()) {
                st = state(t);
                g = lgrid[t];
                if (!gvalid[t]) {
                    p2g_into(st, t, g);
                    gvalid[t] = 1;
                }
            } else {
                st = state(t);
                g = grid;
                p2g_into(st, t, g);
            }
            if (K <= 1 && fresh[t + 1] && valid[t + 1]) nxt = layer[t + 1];
            // Layer t+1 in another order (the first layer of the next
            // segment), where the adjoint reads it: its v and C planes (all
            // the adjoint reads of it) permuted into this order in a scratch
            // state (the backward pass uses neither).
            if (nxt && ord_of(t + 1) != o && !cfg.adjoint_recompute &&
                (cfg.adjoint == DMPM_ADJOINT_REFERENCE || !cfg.trust_region)) {
                for (const int pl : {(int)PV, (int)PC})
                    permute_planes(pr, (const char*)nxt + pl * ld * esz, (char*)scratch[0] + pl * ld * esz,
                                   oseg[o].bwdq, n, pl == PV ? 3 : 9, ld, stream);
                nxt = scratch[0];
            }
            Launch Lc = launch(o);
            Lc.st = st;
            Lc.stat = stat_of(o);
            Lc.ctl = ctl[t];
            Lc.grid = g;
            Lc.ggv = gcur;
            Lc.nxt = cfg.adjoint_recompute ? nullptr : nxt;
            Lc.adj1 = carry[carry_cur];
            Lc.adj0 = carry[carry_cur ^ 1];
            // the first layer of a segment, with the previous segment's last
            // layer still to come: its carry goes out in that segment's order
            // (this step's gather runs unfused, see below)
            if (t > stop && ord_of(t - 1) != o) Lc.opos = oseg[o - 1].fwdq;
            void* gout = nullptr;
            if (t == stop) gout = gctl_stop;
            else if (want_gctl[t]) {
                if (!gctl[t]) gctl[t] = dalloc(ctl_bytes());
                gout = gctl[t];
            }
            Lc.gctl = gout;
            if (cfg.deterministic) {
                if (rec) throw Error(DMPM_ERR_STATE, "deterministic mode cannot be batched");
                timed(2, [&] { det_adj_scatter(pr, Lc, det_buf(), stream); });
            } else if (!scattered) {
                op_zero(gcur, 4 * gld);
                op_kernel(OP_ADJ_SCATTER, Lc);
            }
            // Fuse this step's gather with the scatter of step t-1 when layer
            // t-1 is resident (stored, or in the current segment) and layer t
            // is its forward step as it stands, and the adjoint reads C' from
            // the next layer.
            const void* prev = t - 1 >= stop ? resident_state(t - 1) : nullptr;
            const bool fuse = prev && !cfg.deterministic && !cfg.adjoint_recompute && fresh[t] &&
                              (cfg.adjoint == DMPM_ADJOINT_REFERENCE || !cfg.trust_region) && ord_of(t - 1) == o;
            if (fuse) {
                Launch Lp = launch(o);
                Lp.st = prev;
                Lp.stat = stat_of(o);
                Lp.ctl = ctl[t - 1];
                Lp.ggv = gnext;
                Lp.nxt = st;
                if (!gnext_clear) op_zero(gnext, 4 * gld);
                // the next step fuses only if it has a previous step
                gnext_clear = t - 2 >= stop;
                if (gnext_clear) {
                    Lc.zero = gfree;
                    Lc.zero_bytes = 4 * gld * (long long)esz;
                }
                op_adj_fused(Lc, Lp);
                void* read = gcur;
                gcur = gnext;
                gnext = gfree;
                gfree = read;
                scattered = true;
            } else {
                op_kernel(OP_ADJ_GATHER, Lc);
                scattered = false;
                gnext_clear = false;
            }
            if (t == stop && want_gctl[t]) {
                if (!gctl[t]) gctl[t] = dalloc(ctl_bytes());
                op_copy(gctl[t], gctl_stop, 9 * ld);
            }
            carry_cur ^= 1;
            carry_layer = t;
        }
        if (!capturing && !rec) HIPCHK(hipEventRecord(ev[3], stream));
        if (stop == L - 1) {
            // no step: dL/d dFc of the last layer is dL/dF
            op_copy(gctl_stop, (char*)carry[carry_cur] + PF * ld * esz, 9 * ld);
        }
    }

    // ---- segment orders ----
    // buf (nc planes) := buf permuted by q (slot j <- index q[j]), through tmp
    // (a state-sized buffer): whole states are swapped in, others copied back.
    void permute_buf(void*& buf, int nc, const int* q, void*& tmp) {
        if (!buf) return;
        permute_planes(pr, buf, tmp, q, n, nc, ld, stream);
        if (nc == NSTATE) std::swap(buf, tmp);
        else copy_elems(pr, buf, tmp, (long long)nc * ld, stream);
    }
    // Every per-particle array that belongs to layer t (its state when
    // stored, control, Adam state, control gradient), permuted by q.
    void permute_layer(int t, const int* q, void*& tmp) {
        if (t < (int)layer.size()) permute_buf(layer[t], NSTATE, q, tmp);
        for (auto* v : {&ctl, &am, &av, &avmax, &gctl})
            if (t < (int)v->size()) permute_buf((*v)[t], 9, q, tmp);
    }
    // Back to one storage order (order 0) for every layer.
    void collapse_orders() {
        if (oseg.empty()) return;
        void* tmp = dalloc(state_bytes());
        for (int t = 0; t < L; ++t) {
            const int o = ord_of(t);
            if (o > 0) permute_layer(t, oseg[o - 1].zq, tmp);
        }
        const int ol = ord_of(L - 1);
        if (ol > 0) {
            permute_buf(adj_last, NSTATE, oseg[ol - 1].zq, tmp);
            permute_buf(rg, 12, oseg[ol - 1].zq, tmp);
        }
        const int oc = carry_layer >= 0 ? ord_of(carry_layer) : 0;
        if (oc > 0) {
            for (auto& c : carry) permute_buf(c, NSTATE, oseg[oc - 1].zq, tmp);
            permute_buf(gctl_stop, 9, oseg[oc - 1].zq, tmp);
        }
        dfree(tmp);
        for (auto& os : oseg) {
            dfree((void*&)os.perm);
            dfree((void*&)os.bst);
            dfree(os.stat);
            dfree((void*&)os.fwdq);
            dfree((void*&)os.bwdq);
            dfree((void*&)os.zq);
        }
        oseg.clear();
        ob.clear();
        dfree(ctl_x);
        seg_base = -1;
    }

    // adj_last[plane_off..] += src (internal-order planes)
    void accumulate_planes(const void* src, int plane_off, int nc) {
        // src holds internal-order planes; add them plane by plane (identity permutation)
        for (int c = 0; c < nc; ++c)
            op_accum((char*)adj_last + (size_t)(plane_off + c) * ld * esz, (const char*)src + (size_t)c * ld * esz, n);
    }

    // Upload a caller array (n x nc, caller order) into a device staging
    // buffer of the caller's dtype; returns a device pointer valid until the
    // next call.
    std::vector<void*> staging;
    const void* stage(const void* src, size_t bytes, int on_device) {
        if (on_device || !src) return src;
        void* d = dalloc(bytes);
        staging.push_back(d);
        HIPCHK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream));
        return d;
    }
    void unstage() {
        if (staging.empty()) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& p : staging) dfree(p);
        staging.clear();
    }
};

namespace {

void check_sim(const dmpm_sim* s) {
    if (!s) throw Error(DMPM_ERR_ARG, "null sim");
}

Phys phys_of(const dmpm_config& c) {
    Phys p{};
    for (int i = 0; i < 3; ++i) {
        p.dims[i] = c.dims[i];
        p.grid_min[i] = c.grid_min[i];
        p.f_ext[i] = c.f_ext[i];
    }
    p.dx = c.dx;
    p.dt = c.dt;
    p.drag = c.drag;
    p.smoothing = c.smoothing;
    p.zeta = c.zeta;
    p.jmin = c.jmin;
    p.order = c.order;
    p.model = c.model;
    p.trust_region = c.trust_region;
    p.wall_cells = c.wall_cells;
    p.adjoint = c.adjoint;
    return p;
}

void validate(const dmpm_config& c) {
    for (int i = 0; i < 3; ++i)
        if (c.dims[i] <= 0) throw Error(DMPM_ERR_ARG, "grid dims must be positive");
    if (!(c.dx > 0)) throw Error(DMPM_ERR_ARG, "dx must be positive");
    if (c.order != 2 && c.order != 3) throw Error(DMPM_ERR_ARG, "order must be 2 or 3");
    if (c.model != 0 && c.model != 1) throw Error(DMPM_ERR_ARG, "model must be 0 (FCR) or 1 (Neo-Hookean)");
    if (c.adjoint != 0 && c.adjoint != 1) throw Error(DMPM_ERR_ARG, "adjoint must be 0 or 1");
    if (c.precision != 0 && c.precision != 1) throw Error(DMPM_ERR_ARG, "precision must be 0 or 1");
    if (c.checkpoint_interval < 0) throw Error(DMPM_ERR_ARG, "checkpoint_interval must be >= 0");
    if (c.block != 0 && (c.block % 64 != 0 || c.block > 256))
        throw Error(DMPM_ERR_ARG, "block must be a multiple of 64, at most 256");
    if (c.lds_bytes < 0 || c.lds_bytes > 64 * 1024) throw Error(DMPM_ERR_ARG, "lds_bytes must be in [0, 65536]");
}

size_t dsize(int dtype) { return dtype == DMPM_DTYPE_F64 ? 8 : 4; }

double host_get(const void* p, int dtype, size_t i) {
    return dtype == DMPM_DTYPE_F64 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

// Sum of partials -> host double(s) (synchronises).
void reduce_to_host(dmpm_sim* s, int nblocks, int stride, double* out) {
    reduce_partials(s->partial, nblocks, stride, s->red, s->stream);
    HIPCHK(hipMemcpyAsync(out, s->red, stride * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
}

int red_blocks_for(long long count) {
    long long nb = (count + 255) / 256;
    return (int)std::max(1LL, std::min(nb, 1024LL));
}

double end_layer_mass_loss(dmpm_sim* s) {
    const void* st = s->state(s->L - 1);
    const int o = s->ord_of(s->L - 1);
    Launch Lc = s->launch(o);
    Lc.st = st;
    Lc.stat = s->stat_of(o);
    Lc.grid = s->mgrid;
    if (s->cfg.deterministic) {
        det_p2g(s->pr, Lc, true, s->det_buf(), s->stream);
    } else {
        s->zero(s->mgrid, s->gld);
        p2g(s->pr, Lc, true, s->stream);
    }
    const int nb = red_blocks_for(s->G);
    mass_sums(s->pr, s->mgrid, s->target_m, s->G, s->partial, nb, s->stream);
    reduce_partials(s->partial, nb, 2, s->red, s->stream);
    mass_loss_nodes(s->pr, s->mgrid, s->target_m, s->G, s->red, nb, s->dldm, s->partial + 2 * 1024, nb, s->stream);
    reduce_partials(s->partial + 2 * 1024, nb, 1, s->red + 2, s->stream);
    Lc.adj0 = s->adj_last;
    mass_loss_particles(s->pr, Lc, s->dldm, s->target_m, s->stream);
    double loss = 0;
    HIPCHK(hipMemcpyAsync(&loss, s->red + 2, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return loss;
}

double control_norm(dmpm_sim* s) {
    const int nb = red_blocks_for(s->n);
    sqnorm9(s->pr, s->gctl_stop, s->n, s->ld, s->partial, nb, s->stream);
    double sq = 0;
    reduce_to_host(s, nb, 1, &sq);
    return std::sqrt(sq);
}

void ensure_adam(dmpm_sim* s, int t) {
    for (auto* v : {&s->am, &s->av, &s->avmax}) {
        if (!(*v)[t]) {
            (*v)[t] = s->dalloc(s->ctl_bytes());
            s->zero((*v)[t], 9 * s->ld);
        }
    }
    if (!s->ctl[t]) {
        s->ctl[t] = s->dalloc(s->ctl_bytes());
        s->zero(s->ctl[t], 9 * s->ld);
    }
}

void descend_adam(dmpm_sim* s, int t, double alpha, double grad_norm, double b1, double b2, double eps, int ts) {
    ensure_adam(s, t);
    s->grids_stale(t);
    // gradient_norm as passed by the caller: Descend_Adam only uses 1/max(gn,1e-12)
    double sq = grad_norm * grad_norm;
    HIPCHK(hipMemcpyAsync(s->red + 4, &sq, sizeof(double), hipMemcpyHostToDevice, s->stream));
    adam(s->pr, s->gctl_stop, s->ctl[t], s->am[t], s->av[t], s->avmax[t], s->n, s->ld, s->red + 4, alpha, b1, b2,
         eps, ts, s->stream);
    HIPCHK(hipStreamSynchronize(s->stream));
}

struct Backup {
    void* buf = nullptr;
};

// OptimizeSingleTimestep's line search state backup (CompGraph.cpp:384-391).
void backup_layer(dmpm_sim* s, int t, void* buf) {
    const long long e = 9 * s->ld;
    copy_elems(s->pr, (char*)buf + 0 * e * s->esz, s->ctl[t], e, s->stream);
    copy_elems(s->pr, (char*)buf + 1 * e * s->esz, s->am[t], e, s->stream);
    copy_elems(s->pr, (char*)buf + 2 * e * s->esz, s->av[t], e, s->stream);
    copy_elems(s->pr, (char*)buf + 3 * e * s->esz, s->avmax[t], e, s->stream);
}
void restore_layer(dmpm_sim* s, int t, const void* buf) {
    const long long e = 9 * s->ld;
    s->grids_stale(t);
    copy_elems(s->pr, s->ctl[t], (const char*)buf + 0 * e * s->esz, e, s->stream);
    copy_elems(s->pr, s->am[t], (const char*)buf + 1 * e * s->esz, e, s->stream);
    copy_elems(s->pr, s->av[t], (const char*)buf + 2 * e * s->esz, e, s->stream);
    copy_elems(s->pr, s->avmax[t], (const char*)buf + 3 * e * s->esz, e, s->stream);
}

// dmpm_p2g_mass / dmpm_calculate_volumes: the particles' masses P2G'd into
// grid_m (added to what it holds) and, when vol is non-null, their volumes
// from the summed node masses, computed in the precision of dtype with
// device memory for the planes (and, for host arrays, staging) allocated
// for the call.
void mass_volume_call(const dmpm_config* cfg, int n, const void* x, const void* m, void* vol, void* grid_m,
                      int dtype, int on_device) {
    if (!cfg || !x || !m || !grid_m) throw Error(DMPM_ERR_ARG, "null argument");
    if (n < 0) throw Error(DMPM_ERR_ARG, "n must be non-negative");
    validate(*cfg);
    const Prec pr = dtype == DMPM_DTYPE_F64 ? PREC_F64 : PREC_F32;
    const size_t es = dsize(dtype);
    const long long ld = round_up(std::max(n, 1), 64);
    const long long G = (long long)cfg->dims[0] * cfg->dims[1] * cfg->dims[2];
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<void*> bufs;
    auto dalloc = [&](size_t bytes) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(DMPM_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
        bufs.push_back(p);
        return p;
    };
    auto release = [&] {
        (void)hipStreamSynchronize(st);
        for (void* p : bufs) (void)hipFree(p);
        (void)hipStreamDestroy(st);
    };
    try {
        void* planes = dalloc((size_t)4 * ld * es);
        const void *dx = x, *dm = m;
        void *dg = grid_m, *dv = vol;
        if (!on_device) {
            void* hx = dalloc((size_t)n * 3 * es);
            void* hm = dalloc((size_t)n * es);
            dg = dalloc((size_t)G * es);
            if (vol) dv = dalloc((size_t)n * es);
            HIPCHK(hipMemcpyAsync(hx, x, (size_t)n * 3 * es, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(hm, m, (size_t)n * es, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(dg, grid_m, (size_t)G * es, hipMemcpyHostToDevice, st));
            dx = hx;
            dm = hm;
        }
        mass_volume_pass(pr, *cfg, n, ld, dx, dm, planes, dg, dv, st);
        if (!on_device) {
            HIPCHK(hipMemcpyAsync(grid_m, dg, (size_t)G * es, hipMemcpyDeviceToHost, st));
            if (vol) HIPCHK(hipMemcpyAsync(vol, dv, (size_t)n * es, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
    } catch (...) {
        release();
        throw;
    }
    release();
}

}  // namespace

// Shared with the allocation-free entry point (dmpm_noalloc.cpp).
void dmpm::mass_volume_pass(Prec pr, const dmpm_config& cfg, int n, long long ld, const void* x, const void* m,
                            void* planes, void* grid_m, void* vol, hipStream_t st) {
    if (n <= 0) return;
    const size_t es = pr == PREC_F64 ? 8 : 4;
    char* P = (char*)planes;
    import_planes(pr, x, pr == PREC_F64, P, nullptr, n, 3, ld, st);
    import_planes(pr, m, pr == PREC_F64, P + 3 * ld * es, nullptr, n, 1, ld, st);
    Launch Lc{};
    Lc.ph = phys_of(cfg);
    Lc.n = n;
    Lc.ld = ld;
    Lc.gld = round_up((long long)cfg.dims[0] * cfg.dims[1] * cfg.dims[2], 64);
    Lc.st = P;
    Lc.stat = P + 3 * ld * es;  // the statics' first plane is m
    Lc.grid = grid_m;           // only the m plane is touched by a mass-only P2G
    Lc.lds_bytes = (pr == PREC_F32 ? 30 : 48) * 1024;
    Lc.block = 256;
    p2g(pr, Lc, true, st);
    if (vol) volumes(pr, Lc, vol, st);
}

// ============================================================================
extern "C" {

void dmpm_config_init(dmpm_config* c, int nx, int ny, int nz, double dx, const double grid_min[3]) {
    std::memset(c, 0, sizeof(*c));
    c->dims[0] = nx;
    c->dims[1] = ny;
    c->dims[2] = nz;
    c->dx = dx;
    for (int i = 0; i < 3; ++i) c->grid_min[i] = grid_min ? grid_min[i] : 0.0;
    c->dt = 1.0 / 120.0;
    c->drag = 0.5;
    c->smoothing = 0.1;
    c->zeta = 0.25;
    c->jmin = 0.6;
    c->order = 3;
    c->model = 0;
    c->trust_region = 1;
    c->wall_cells = 0;
    c->adjoint = 1;
    c->precision = DMPM_PREC_F32;
    c->checkpoint_interval = 1;
}

const char* dmpm_last_error(void) { return g_err.c_str(); }

int dmpm_get_config(const dmpm_sim* s, dmpm_config* out) {
    return guard([&] {
        check_sim(s);
        if (!out) throw Error(DMPM_ERR_ARG, "null output");
        *out = s->cfg;
    });
}
const char* dmpm_version(void) { return "dmpm 0.1 (gfx950)"; }

int dmpm_create(const dmpm_config* cfg, dmpm_sim** out) {
    return guard([&] {
        if (!cfg || !out) throw Error(DMPM_ERR_ARG, "null argument");
        validate(*cfg);
        auto s = std::make_unique<dmpm_sim>();
        s->cfg = *cfg;
        s->ph = phys_of(*cfg);
        s->pr = cfg->precision == DMPM_PREC_F64 ? PREC_F64 : PREC_F32;
        s->esz = s->pr == PREC_F64 ? 8 : 4;
        s->K = std::max(1, cfg->checkpoint_interval);
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
        for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
        s->G = (long long)cfg->dims[0] * cfg->dims[1] * cfg->dims[2];
        s->gld = round_up(s->G, 64);
        // the kernels address a grid with 32-bit byte offsets
        if ((unsigned long long)4 * s->gld * s->esz >= (1ull << 32))
            throw Error(DMPM_ERR_ARG, "grid too large: its four values per node must stay below 4 GiB");
        s->grid = s->dalloc(4 * s->gld * s->esz);
        s->grid2 = s->dalloc(4 * s->gld * s->esz);
        s->ggv = s->dalloc(4 * s->gld * s->esz);
        s->ggv2 = s->dalloc(4 * s->gld * s->esz);
        s->ggv3 = s->dalloc(4 * s->gld * s->esz);
        s->target_m = s->dalloc(s->gld * s->esz);
        s->mgrid = s->dalloc(s->gld * s->esz);
        s->dldm = s->dalloc(s->gld * s->esz);
        s->zero(s->target_m, s->gld);
        s->partial = (double*)s->dalloc(4 * 1024 * sizeof(double));
        s->red = (double*)s->dalloc(8 * sizeof(double));
        s->rgn = (double*)s->dalloc(2 * sizeof(double));
        s->wstat = (unsigned long long*)s->dalloc(2 * sizeof(unsigned long long));
        s->zero(s->wstat, 4);  // 16 bytes
        *out = s.release();
    });
}

int dmpm_destroy(dmpm_sim* sim) {
    return guard([&] { delete sim; });
}

int dmpm_set_stream(dmpm_sim* s, void* stream) {
    return guard([&] {
        check_sim(s);
        HIPCHK(hipStreamSynchronize(s->stream));
        if (stream) {
            if (s->own_stream) HIPCHK(hipStreamDestroy(s->stream));
            s->stream = (hipStream_t)stream;
            s->own_stream = false;
        }
    });
}

int dmpm_synchronize(dmpm_sim* s) {
    return guard([&] {
        check_sim(s);
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

}  // extern "C"

namespace {
// Storage order and thread blocks of n particles (order_particles_device,
// dmpm_kernels.h) from device positions x: caller (n,3) layout (soa = 0) of
// dtype f64/f32, or three planes of stride ld of the sim's precision.
// Returns the permutation (device, n ints: source index of each stored
// particle) and the block table (device, nblk + 1 starts; null when blocks
// are not cut), both allocated on the sim.
struct Order {
    int* perm = nullptr;
    int* bst = nullptr;
    int nblk = 0;
};
Order order_on_device(dmpm_sim* s, const void* x, int x_f64, int soa, long long ld, int n,
                      const void* const* xcut = nullptr, int ncut = 0) {
    Order o;
    const size_t sb = order_scratch_bytes(n, ncut);
    void* scr = s->dalloc(sb);
    int* starts = (int*)s->dalloc(((size_t)n + 1) * sizeof(int));
    o.perm = (int*)s->dalloc((size_t)n * sizeof(int));
    try {
        o.nblk = order_particles_device(s->cfg.order, s->cfg.block, s->cfg.split_nodes, s->cfg.grid_min, s->cfg.dx, x,
                                        x_f64, soa, ld, n, o.perm, starts, scr, sb, s->stream, xcut, ncut);
    } catch (const std::exception& e) {
        s->dfree(scr);
        s->dfree((void*&)starts);
        throw Error(DMPM_ERR_HIP, e.what());
    }
    if (o.nblk > 0) {
        o.bst = (int*)s->dalloc(((size_t)o.nblk + 1) * sizeof(int));
        HIPCHK(hipMemcpyAsync(o.bst, starts, ((size_t)o.nblk + 1) * sizeof(int), hipMemcpyDeviceToDevice,
                              s->stream));
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    s->dfree(scr);
    s->dfree((void*&)starts);
    return o;
}

}  // namespace

extern "C" {

int dmpm_set_particles(dmpm_sim* s, int n, const void* x, const void* v, const void* F, const void* C,
                       const void* dFc, const void* m, const void* vol, const void* lam, const void* mu, int dtype,
                       int on_device) {
    return guard([&] {
        check_sim(s);
        s->refuse_if_captured("dmpm_set_particles");
        if (n <= 0) throw Error(DMPM_ERR_ARG, "n must be positive");
        if (!x || !v || !F || !C || !m || !vol || !lam || !mu) throw Error(DMPM_ERR_ARG, "null particle array");
        // the kernels address a particle array with 32-bit byte offsets
        if ((unsigned long long)NSTATE * round_up(n, 64) * s->esz >= (1ull << 32))
            throw Error(DMPM_ERR_ARG, "too many particles for one sim: a layer must stay below 4 GiB (" +
                                          std::to_string((1ull << 32) / (NSTATE * s->esz)) + " particles)");
        const size_t es = dsize(dtype);
        s->collapse_orders();
        // storage order from the positions on the device (host positions
        // are uploaded first; device ones are never copied to the host)
        const void* xd = s->stage(x, (size_t)n * 3 * es, on_device);
        const Order ord = order_on_device(s, xd, dtype == DMPM_DTYPE_F64, 0, 0, n);

        // (re)allocate particle storage
        for (auto& p : s->layer) s->dfree(p);
        for (auto& p : s->ctl) s->dfree(p);
        for (auto* v2 : {&s->am, &s->av, &s->avmax, &s->gctl})
            for (auto& p : *v2) s->dfree(p);
        for (size_t j = 1; j < s->seg.size(); ++j) s->dfree(s->seg[j]);
        for (auto& g : s->seg_grid) s->dfree(g);
        for (auto& p : s->scratch) s->dfree(p);
        s->seg.clear();
        s->seg_grid.clear();
        s->dfree(s->stat);
        s->dfree((void*&)s->perm);
        s->dfree((void*&)s->bst);
        s->nblk = 0;
        for (auto& p : s->carry) s->dfree(p);
        s->dfree(s->adj_last);
        s->dfree(s->gctl_stop);
        s->dfree(s->rg);
        s->dfree(s->det);
        s->has_rg = false;

        if (round_up(n, 64) != s->ld) s->release_pool();  // blocks of the old size fit nothing now
        s->n = n;
        s->ld = round_up(n, 64);
        s->L = 1;
        s->layer.assign(1, nullptr);
        s->valid.assign(1, 1);
        s->pend.assign(1, 0);
        s->fresh.assign(1, 0);
        s->ctl.assign(1, nullptr);
        s->am.assign(1, nullptr);
        s->av.assign(1, nullptr);
        s->avmax.assign(1, nullptr);
        s->gctl.assign(1, nullptr);
        s->want_gctl.assign(1, 0);
        s->nblk = ord.nblk;
        s->bst = ord.bst;
        s->perm = ord.perm;
        s->layer[0] = s->dalloc(s->state_bytes());
        s->stat = s->dalloc((size_t)NSTAT * s->ld * s->esz);
        for (auto& p : s->carry) p = s->dalloc(s->state_bytes());
        s->adj_last = s->dalloc(s->state_bytes());
        s->zero(s->adj_last, NSTATE * s->ld);
        s->gctl_stop = s->dalloc(s->ctl_bytes());
        s->zero(s->gctl_stop, 9 * s->ld);
        s->zero(s->layer[0], NSTATE * s->ld);
        s->zero(s->stat, NSTAT * s->ld);
        const int f64 = dtype == DMPM_DTYPE_F64;
        auto imp = [&](const void* src, int nc, void* dst) {
            const void* d = s->stage(src, (size_t)n * nc * es, on_device);
            import_planes(s->pr, d, f64, dst, s->perm, n, nc, s->ld, s->stream);
        };
        char* L0 = (char*)s->layer[0];
        imp(x, 3, L0 + PX * s->ld * s->esz);
        imp(v, 3, L0 + PV * s->ld * s->esz);
        imp(F, 9, L0 + PF * s->ld * s->esz);
        imp(C, 9, L0 + PC * s->ld * s->esz);
        char* S = (char*)s->stat;
        imp(m, 1, S + SM * s->ld * s->esz);
        imp(vol, 1, S + SVOL * s->ld * s->esz);
        imp(lam, 1, S + SLAM * s->ld * s->esz);
        imp(mu, 1, S + SMU * s->ld * s->esz);
        if (dFc) {
            s->ctl[0] = s->dalloc(s->ctl_bytes());
            imp(dFc, 9, s->ctl[0]);
        }
        s->unstage();
        s->alloc_layers();
    });
}

int dmpm_num_particles(const dmpm_sim* s, int* n) {
    return guard([&] {
        check_sim(s);
        *n = s->n;
    });
}

int dmpm_set_target_grid_mass(dmpm_sim* s, const void* m, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (!m) throw Error(DMPM_ERR_ARG, "null target mass");
        const void* d = s->stage(m, (size_t)s->G * dsize(dtype), on_device);
        import_planes(s->pr, d, dtype == DMPM_DTYPE_F64, s->target_m, nullptr, (int)s->G, 1, s->gld, s->stream);
        s->unstage();
    });
}

int dmpm_set_sim_params(dmpm_sim* s, double dt, double drag, double smoothing, const double f_ext[3]) {
    return guard([&] {
        check_sim(s);
        // captured launches carry dt, drag, smoothing and f_ext by value
        s->refuse_if_captured("dmpm_set_sim_params");
        s->cfg.dt = dt;
        s->cfg.drag = drag;
        s->cfg.smoothing = smoothing;
        if (f_ext)
            for (int i = 0; i < 3; ++i) s->cfg.f_ext[i] = f_ext[i];
        s->ph = phys_of(s->cfg);
        s->grids_stale(-1);
    });
}

int dmpm_set_up_comp_graph(dmpm_sim* s, int num_layers) {
    return guard([&] {
        check_sim(s);
        s->refuse_if_captured("dmpm_set_up_comp_graph");
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        if (num_layers <= 0) throw Error(DMPM_ERR_ARG, "num_layers must be positive");
        s->collapse_orders();
        for (int t = 1; t < (int)s->ctl.size(); ++t) {
            s->dfree(s->ctl[t]);
            s->dfree(s->am[t]);
            s->dfree(s->av[t]);
            s->dfree(s->avmax[t]);
            s->dfree(s->gctl[t]);
        }
        s->L = num_layers;
        s->ctl.resize(s->L, nullptr);
        s->am.resize(s->L, nullptr);
        s->av.resize(s->L, nullptr);
        s->avmax.resize(s->L, nullptr);
        s->gctl.resize(s->L, nullptr);
        s->want_gctl.resize(s->L, 0);
        s->alloc_layers();
        // every layer is a copy of layer 0, controls and Adam state included
        // (the state copies are made when a layer is first read, unless the
        // forward pass overwrites it first)
        for (int t = 1; t < s->L; ++t) {
            if (s->layer[t]) {
                s->pend[t] = 1;
                s->valid[t] = 1;
            }
            auto dup = [&](std::vector<void*>& v) {
                if (v[0]) {
                    v[t] = s->dalloc(s->ctl_bytes());
                    copy_elems(s->pr, v[t], v[0], 9 * s->ld, s->stream);
                }
            };
            dup(s->ctl);
            dup(s->am);
            dup(s->av);
            dup(s->avmax);
        }
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

int dmpm_get_num_layers(const dmpm_sim* s, int* num_layers) {
    return guard([&] {
        check_sim(s);
        *num_layers = s->L;
    });
}

int dmpm_compute_forward_pass(dmpm_sim* s, int start_layer) {
    return guard([&] {
        check_sim(s);
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        s->forward(start_layer);
        HIPCHK(hipGetLastError());
    });
}

int dmpm_rollout_forward(dmpm_sim* s, int start_layer) {
    return guard([&] {
        check_sim(s);
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        s->forward(start_layer, false);
        HIPCHK(hipGetLastError());
    });
}

int dmpm_end_layer_mass_loss(dmpm_sim* s, double* loss) {
    return guard([&] {
        check_sim(s);
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        double l = end_layer_mass_loss(s);
        if (loss) *loss = l;
    });
}

int dmpm_compute_backward_pass(dmpm_sim* s, int control_layer) {
    return guard([&] {
        check_sim(s);
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        s->backward(control_layer);
        HIPCHK(hipGetLastError());
    });
}

int dmpm_get_layer(dmpm_sim* s, int layer, int field, void* out, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        if (!out) throw Error(DMPM_ERR_ARG, "null output");
        // the layer's storage order (segment orders) and its statics
        const int* pm = s->perm_of(s->ord_of(layer));
        void* const sto = s->stat_of(s->ord_of(layer));
        int nc = 0, plane = 0;
        const void* base = nullptr;
        const size_t es = dsize(dtype);
        switch (field) {
            case DMPM_FIELD_X: nc = 3; plane = PX; break;
            case DMPM_FIELD_V: nc = 3; plane = PV; break;
            case DMPM_FIELD_F: nc = 9; plane = PF; break;
            case DMPM_FIELD_C: nc = 9; plane = PC; break;
            case DMPM_FIELD_DFC: nc = 9; break;
            case DMPM_FIELD_FTOTAL: nc = 9; break;
            case DMPM_FIELD_M: nc = 1; plane = SM; base = sto; break;
            case DMPM_FIELD_VOL: nc = 1; plane = SVOL; base = sto; break;
            case DMPM_FIELD_LAM: nc = 1; plane = SLAM; base = sto; break;
            case DMPM_FIELD_MU: nc = 1; plane = SMU; base = sto; break;
            case DMPM_FIELD_P: nc = 9; break;
            case DMPM_FIELD_ADAM_M: nc = 9; base = s->am[layer]; break;
            case DMPM_FIELD_ADAM_V: nc = 9; base = s->av[layer]; break;
            case DMPM_FIELD_ADAM_VMAX: nc = 9; base = s->avmax[layer]; break;
            default: throw Error(DMPM_ERR_ARG, "unknown field");
        }
        const size_t bytes = (size_t)s->n * nc * es;
        void* dst = on_device ? out : s->dalloc(bytes);
        const int f64 = dtype == DMPM_DTYPE_F64;
        if (field == DMPM_FIELD_ADAM_M || field == DMPM_FIELD_ADAM_V || field == DMPM_FIELD_ADAM_VMAX) {
            if (base) export_planes(s->pr, base, dst, f64, pm, s->n, 9, s->ld, s->stream);
            else fill_zero(f64 ? PREC_F64 : PREC_F32, dst, (long long)s->n * 9, s->stream);
        } else if (field == DMPM_FIELD_P) {
            export_stress(s->pr, s->cfg.model, s->state(layer), s->ctl[layer], sto, dst, f64, pm, s->n,
                          s->ld, s->stream);
        } else if (field == DMPM_FIELD_DFC) {
            if (s->ctl[layer]) export_planes(s->pr, s->ctl[layer], dst, f64, pm, s->n, 9, s->ld, s->stream);
            else fill_zero(dtype == DMPM_DTYPE_F64 ? PREC_F64 : PREC_F32, dst, (long long)s->n * 9, s->stream);
        } else if (field == DMPM_FIELD_FTOTAL) {
            export_F_total(s->pr, s->state(layer), s->ctl[layer], dst, f64, pm, s->n, s->ld, s->stream);
        } else {
            if (!base) base = s->state(layer);
            export_planes(s->pr, (const char*)base + plane * s->ld * s->esz, dst, f64, pm, s->n, nc, s->ld,
                          s->stream);
        }
        if (!on_device) {
            HIPCHK(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            s->dfree(dst);
        }
    });
}

int dmpm_get_layer_grad(dmpm_sim* s, int layer, int field, void* out, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        const void* src = nullptr;
        int nc = 0, plane = 0;
        switch (field) {
            case DMPM_FIELD_X: nc = 3; plane = PX; break;
            case DMPM_FIELD_V: nc = 3; plane = PV; break;
            case DMPM_FIELD_F: nc = 9; plane = PF; break;
            case DMPM_FIELD_C: nc = 9; plane = PC; break;
            case DMPM_FIELD_DFC: nc = 9; break;
            default: throw Error(DMPM_ERR_ARG, "unknown gradient field");
        }
        if (field == DMPM_FIELD_DFC) {
            if (layer != s->carry_layer) throw Error(DMPM_ERR_STATE, "no control gradient for this layer");
            src = s->gctl_stop;
        } else if (layer == s->carry_layer) {
            src = (const char*)s->carry[s->carry_cur] + plane * s->ld * s->esz;
        } else if (layer == s->L - 1) {
            src = (const char*)s->adj_last + plane * s->ld * s->esz;
        } else {
            throw Error(DMPM_ERR_STATE, "gradients are kept for the control layer and the last layer only");
        }
        const size_t bytes = (size_t)s->n * nc * dsize(dtype);
        void* dst = on_device ? out : s->dalloc(bytes);
        export_planes(s->pr, src, dst, dtype == DMPM_DTYPE_F64, s->perm_of(s->ord_of(layer)), s->n, nc, s->ld, s->stream);
        if (!on_device) {
            HIPCHK(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            s->dfree(dst);
        }
    });
}

int dmpm_accumulate_last_layer_grad(dmpm_sim* s, const void* dLdF, const void* dLdx, const void* dLdv,
                                    const void* dLdC, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        const size_t es = dsize(dtype);
        const int f64 = dtype == DMPM_DTYPE_F64;
        auto acc = [&](const void* src, int off, int nc) {
            if (!src) return;
            const void* d = s->stage(src, (size_t)s->n * nc * es, on_device);
            accumulate_carry(s->pr, d, f64, s->adj_last, off, nc, s->perm_of(s->ord_of(s->L - 1)), s->n, s->ld, s->stream);
        };
        acc(dLdF, PF, 9);
        acc(dLdx, PX, 3);
        acc(dLdv, PV, 3);
        acc(dLdC, PC, 9);
        s->unstage();
    });
}

int dmpm_set_last_layer_grad(dmpm_sim* s, const void* dLdx, const void* dLdv, const void* dLdF, const void* dLdC,
                             int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        s->zero(s->adj_last, NSTATE * s->ld);
        const size_t es = dsize(dtype);
        const int f64 = dtype == DMPM_DTYPE_F64;
        auto put = [&](const void* src, int off, int nc) {
            if (!src) return;
            const void* d = s->stage(src, (size_t)s->n * nc * es, on_device);
            import_planes(s->pr, d, f64, (char*)s->adj_last + off * s->ld * s->esz, s->perm_of(s->ord_of(s->L - 1)), s->n, nc, s->ld,
                          s->stream);
        };
        put(dLdx, PX, 3);
        put(dLdv, PV, 3);
        put(dLdF, PF, 9);
        put(dLdC, PC, 9);
        s->unstage();
    });
}

int dmpm_set_render_gradients(dmpm_sim* s, int n, const void* dLdF, const void* dLdx, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (!dLdF || !dLdx || n < 0) throw Error(DMPM_ERR_ARG, "bad render gradients");
        const size_t es = dsize(dtype);
        const int f64 = dtype == DMPM_DTYPE_F64;
        const Prec dp = f64 ? PREC_F64 : PREC_F32;
        // A capture injects from the buffer it recorded: only an update in place is possible.
        if (s->captured() && (!s->rg || n != s->n))
            s->refuse_if_captured("dmpm_set_render_gradients without a buffer of this size at capture");
        const void* dF = s->stage(dLdF, (size_t)n * 9 * es, on_device);
        const void* dX = s->stage(dLdx, (size_t)n * 3 * es, on_device);
        // The norms of the caller's data (bind.cpp:413-422), summed on the
        // device in a fixed order; read back only by dmpm_render_gradient_info.
        const int nb = red_blocks_for(std::max<long long>((long long)n * 9, 1));
        sqnorm_planes(dp, dF, n * 9, 1, 0, s->partial, nb, s->stream);
        reduce_partials(s->partial, nb, 1, s->rgn, s->stream);
        sqnorm_planes(dp, dX, n * 3, 1, 0, s->partial + 1024, nb, s->stream);
        reduce_partials(s->partial + 1024, nb, 1, s->rgn + 1, s->stream);
        s->rg_norms_host = false;
        s->rg_n = n;
        s->has_rg = true;
        if (n == s->n) {
            if (!s->rg) s->rg = s->dalloc((size_t)12 * s->ld * s->esz);
            import_planes(s->pr, dF, f64, s->rg, s->perm_of(s->ord_of(s->L - 1)), n, 9, s->ld, s->stream);
            import_planes(s->pr, dX, f64, (char*)s->rg + 9 * s->ld * s->esz, s->perm_of(s->ord_of(s->L - 1)), n, 3, s->ld, s->stream);
        } else {
            // stored for the info call only: the backward pass skips a
            // mismatched set (CompGraph.cpp:148-160)
            s->dfree(s->rg);
        }
        s->unstage();
    });
}

int dmpm_clear_render_gradients(dmpm_sim* s) {
    return guard([&] {
        check_sim(s);
        // a capture keeps injecting from its buffer: zeros inject nothing
        if (s->captured() && s->rg) s->zero(s->rg, 12 * s->ld);
        else s->dfree(s->rg);
        s->has_rg = false;
        s->rg_n = 0;
        s->rg_normF = s->rg_normx = 0;
        s->rg_norms_host = true;
    });
}

int dmpm_render_gradient_info(const dmpm_sim* cs, int* has, int* num_points, double* nF, double* nx) {
    return guard([&] {
        check_sim(cs);
        dmpm_sim* s = const_cast<dmpm_sim*>(cs);
        if (s->has_rg && !s->rg_norms_host) {
            double h[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(h, s->rgn, sizeof(h), hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            s->rg_normF = std::sqrt(h[0]);
            s->rg_normx = std::sqrt(h[1]);
            s->rg_norms_host = true;
        }
        if (has) *has = s->has_rg;
        if (num_points) *num_points = s->rg_n;
        if (nF) *nF = s->has_rg ? s->rg_normF : 0.0;
        if (nx) *nx = s->has_rg ? s->rg_normx : 0.0;
    });
}

int dmpm_sort_particles(dmpm_sim* s, int layer) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        if (*s->captures > 0)
            throw Error(DMPM_ERR_STATE, "captured graphs or batched episodes hold this sim's buffers; destroy them "
                                        "before sorting its particles");
        s->collapse_orders();
        const void* st = s->state(layer);
        const long long ld = s->ld;
        // The blocks are cut so that their windows stay within the bound at
        // the first and the last layer of the episode too (where computed):
        // a body that deforms strongly spreads a block cut at one layer over
        // a larger box at another.
        const void* xcut[kOrderCutSets - 1];
        int ncut = 0;
        for (int t : {0, s->L - 1})
            if (t != layer && s->stored(t) && s->valid[t] && !s->pend[t] && s->layer[t] && ncut < kOrderCutSets - 1)
                xcut[ncut++] = (const char*)s->layer[t] + PX * ld * s->esz;
        // q[j]: the stored index of the particle that moves to slot j
        const Order ord = order_on_device(s, (const char*)st + PX * ld * s->esz, s->pr == PREC_F64, 1, ld, s->n, xcut,
                                          ncut);
        int* dq = ord.perm;
        void* tmp = s->dalloc(s->state_bytes());
        // every per-particle array in the internal order: into tmp, then
        // swapped in (whole states) or copied back
        auto reorder = [&](void*& buf, int nc) {
            if (!buf) return;
            permute_planes(s->pr, buf, tmp, dq, s->n, nc, ld, s->stream);
            if (nc == NSTATE) std::swap(buf, tmp);
            else copy_elems(s->pr, buf, tmp, (long long)nc * ld, s->stream);
        };
        for (auto& p : s->layer) reorder(p, NSTATE);
        for (auto& p : s->carry) reorder(p, NSTATE);
        reorder(s->adj_last, NSTATE);
        reorder(s->stat, NSTAT);
        for (auto* v : {&s->ctl, &s->am, &s->av, &s->avmax, &s->gctl})
            for (auto& p : *v) reorder(p, 9);
        reorder(s->gctl_stop, 9);
        reorder(s->rg, 12);
        int* perm2 = (int*)s->dalloc((size_t)s->n * sizeof(int));
        permute_ints(s->perm, perm2, dq, s->n, s->stream);
        HIPCHK(hipStreamSynchronize(s->stream));
        s->dfree((void*&)s->perm);
        s->perm = perm2;
        s->dfree(tmp);
        s->dfree((void*&)dq);
        s->dfree((void*&)s->bst);
        s->nblk = ord.nblk;
        s->bst = ord.bst;
        // recomputed segments were in the old order
        s->seg_base = -1;
    });
}

int dmpm_sort_segments(dmpm_sim* s, int segments) {
    return guard([&] {
        check_sim(s);
        if (segments < 1) throw Error(DMPM_ERR_ARG, "segments must be at least 1");
        if (*s->captures > 0)
            throw Error(DMPM_ERR_STATE, "captured graphs or batched episodes hold this sim's buffers; destroy them "
                                        "before sorting its particles");
        if (s->cfg.deterministic) throw Error(DMPM_ERR_STATE, "segment orders are not used in deterministic mode");
        if (s->n <= 0) throw Error(DMPM_ERR_STATE, "no particles");
        s->collapse_orders();
        s->materialize_all();  // a copy of layer 0 in another order would need its own permutation
        const int L = s->L, K = s->K, n = s->n;
        const long long ld = s->ld;
        // segment boundaries: equal lengths, multiples of the checkpoint
        // interval (a recomputed checkpoint segment stays in one order)
        int P = (L + segments - 1) / segments;
        if (K > 1) P = (P + K - 1) / K * K;
        std::vector<int> b;
        for (int t = 0; t < L; t += P) b.push_back(t);
        b.push_back(L);
        const int S = (int)b.size() - 1;
        // each segment's order: the Hilbert order of its middle layer, its
        // blocks cut so that their windows also fit at the stored layers
        // that bound it
        std::vector<Order> ord(S);
        try {
            for (int q = 0; q < S; ++q) {
                const int mid = (b[q] + b[q + 1] - 1) / 2;
                const void* st = s->state(mid);
                const void* xcut[kOrderCutSets - 1];
                int ncut = 0;
                for (int t : {b[q], std::min(b[q + 1], L - 1)})
                    if (t != mid && s->stored(t) && s->valid[t] && !s->pend[t] && s->layer[t] && ncut < kOrderCutSets - 1)
                        xcut[ncut++] = (const char*)s->layer[t] + PX * ld * s->esz;
                ord[q] = order_on_device(s, (const char*)st + PX * ld * s->esz, s->pr == PREC_F64, 1, ld, n, xcut,
                                         ncut);
            }
        } catch (...) {
            for (auto& o : ord) {
                s->dfree((void*&)o.perm);
                s->dfree((void*&)o.bst);
            }
            throw;
        }
        // ord[q].perm = dq_q: slot j of order q <- index in the current order
        std::vector<int*> inv(S);
        for (int q = 0; q < S; ++q) {
            inv[q] = (int*)s->dalloc((size_t)n * sizeof(int));
            invert_perm(ord[q].perm, inv[q], n, s->stream);
        }
        auto seg_of = [&](int t) {
            int q = 0;
            while (q + 1 < S && t >= b[q + 1]) ++q;
            return q;
        };
        std::vector<dmpm_sim::OrderSeg> os(S);
        for (int q = 0; q < S; ++q) {
            os[q].perm = (int*)s->dalloc((size_t)n * sizeof(int));
            permute_ints(s->perm, os[q].perm, ord[q].perm, n, s->stream);
            os[q].bst = ord[q].bst;
            os[q].nblk = ord[q].nblk;
            os[q].stat = s->dalloc((size_t)NSTAT * ld * s->esz);
            permute_planes(s->pr, s->stat, os[q].stat, ord[q].perm, n, NSTAT, ld, s->stream);
            if (q == 0) continue;
            os[q].fwdq = (int*)s->dalloc((size_t)n * sizeof(int));
            permute_ints(inv[q - 1], os[q].fwdq, ord[q].perm, n, s->stream);
            os[q].bwdq = (int*)s->dalloc((size_t)n * sizeof(int));
            permute_ints(inv[q], os[q].bwdq, ord[q - 1].perm, n, s->stream);
            os[q].zq = (int*)s->dalloc((size_t)n * sizeof(int));
            permute_ints(inv[q], os[q].zq, ord[0].perm, n, s->stream);
        }
        // every per-particle array into the order of the layer it belongs to
        void* tmp = s->dalloc(s->state_bytes());
        for (int t = 0; t < L; ++t) s->permute_layer(t, ord[seg_of(t)].perm, tmp);
        s->permute_buf(s->adj_last, NSTATE, ord[seg_of(L - 1)].perm, tmp);
        s->permute_buf(s->rg, 12, ord[seg_of(L - 1)].perm, tmp);
        const int qc = s->carry_layer >= 0 ? seg_of(s->carry_layer) : 0;
        for (auto& c : s->carry) s->permute_buf(c, NSTATE, ord[qc].perm, tmp);
        s->permute_buf(s->gctl_stop, 9, ord[qc].perm, tmp);
        s->dfree(tmp);
        // install: order 0 in place of the single order, the others after it
        s->dfree((void*&)s->perm);
        s->dfree((void*&)s->bst);
        s->dfree(s->stat);
        s->perm = os[0].perm;
        s->bst = os[0].bst;
        s->nblk = os[0].nblk;
        s->stat = os[0].stat;
        s->oseg.assign(os.begin() + 1, os.end());
        s->ob = b;
        // staging of a boundary layer's control (allocated here, not while a
        // forward pass is being captured)
        if (S > 1) s->ctl_x = s->dalloc(s->ctl_bytes());
        s->seg_base = -1;
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int q = 0; q < S; ++q) {
            s->dfree((void*&)ord[q].perm);
            s->dfree((void*&)inv[q]);
        }
    });
}

int dmpm_promote_last_as_initial(dmpm_sim* s, int /*carry_grid*/) {
    return guard([&] {
        check_sim(s);
        s->refuse_if_captured("dmpm_promote_last_as_initial");
        s->collapse_orders();
        if (s->L <= 0) return;
        const int last = s->L - 1;
        if (last > 0) {
            const void* st = s->state(last);
            copy_elems(s->pr, s->layer[0], st, NSTATE * s->ld, s->stream);
            auto mv = [&](std::vector<void*>& v) {
                s->dfree(v[0]);
                v[0] = v[last];
                v[last] = nullptr;
            };
            mv(s->ctl);
            mv(s->am);
            mv(s->av);
            mv(s->avmax);
        }
        for (int t = 1; t < s->L; ++t) {
            s->dfree(s->ctl[t]);
            s->dfree(s->am[t]);
            s->dfree(s->av[t]);
            s->dfree(s->avmax[t]);
            s->dfree(s->gctl[t]);
        }
        s->L = 1;
        s->ctl.resize(1);
        s->am.resize(1);
        s->av.resize(1);
        s->avmax.resize(1);
        s->gctl.resize(1);
        s->want_gctl.resize(1);
        s->alloc_layers();
        s->carry_layer = -1;
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

int dmpm_set_control(dmpm_sim* s, int layer, const void* dFc, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        if (!dFc) {
            // a capture keeps reading the buffer: zero it in place
            if (s->captured() && s->ctl[layer]) s->zero(s->ctl[layer], 9 * s->ld);
            else s->dfree(s->ctl[layer]);
            s->grids_stale(layer);
            return;
        }
        if (!s->ctl[layer]) {
            // a capture recorded this layer without a control and would ignore a new buffer
            s->refuse_if_captured("dmpm_set_control on a layer that had no control when captured");
            s->ctl[layer] = s->dalloc(s->ctl_bytes());
        }
        const void* d = s->stage(dFc, (size_t)s->n * 9 * dsize(dtype), on_device);
        import_planes(s->pr, d, dtype == DMPM_DTYPE_F64, s->ctl[layer], s->perm_of(s->ord_of(layer)), s->n, 9, s->ld, s->stream);
        s->unstage();
        s->grids_stale(layer);
    });
}

int dmpm_set_adam_state(dmpm_sim* s, int layer, const void* m, const void* v, const void* vmax, int dtype,
                        int on_device) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        if (!m || !v || !vmax) throw Error(DMPM_ERR_ARG, "null Adam state");
        ensure_adam(s, layer);
        const size_t bytes = (size_t)s->n * 9 * dsize(dtype);
        const int f64 = dtype == DMPM_DTYPE_F64;
        const void* src[3] = {m, v, vmax};
        void* dst[3] = {s->am[layer], s->av[layer], s->avmax[layer]};
        for (int q = 0; q < 3; ++q) {
            const void* d = s->stage(src[q], bytes, on_device);
            import_planes(s->pr, d, f64, dst[q], s->perm_of(s->ord_of(layer)), s->n, 9, s->ld, s->stream);
        }
        s->unstage();
    });
}

int dmpm_request_control_grad(dmpm_sim* s, int layer, int on) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        s->want_gctl[layer] = on ? 1 : 0;
        if (!on && !s->captured()) s->dfree(s->gctl[layer]);
    });
}

int dmpm_get_control_grad(dmpm_sim* s, int layer, void* out, int dtype, int on_device) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        const void* src = s->gctl[layer];
        if (!src) {
            if (layer == s->carry_layer) src = s->gctl_stop;
            else throw Error(DMPM_ERR_STATE, "no control gradient recorded for this layer");
        }
        const size_t bytes = (size_t)s->n * 9 * dsize(dtype);
        void* dst = on_device ? out : s->dalloc(bytes);
        export_planes(s->pr, src, dst, dtype == DMPM_DTYPE_F64, s->perm_of(s->ord_of(layer)), s->n, 9, s->ld, s->stream);
        if (!on_device) {
            HIPCHK(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            s->dfree(dst);
        }
    });
}

int dmpm_control_grad_norm(dmpm_sim* s, double* norm) {
    return guard([&] {
        check_sim(s);
        *norm = control_norm(s);
    });
}

int dmpm_descend_adam(dmpm_sim* s, int layer, double alpha, double grad_norm, double beta1, double beta2,
                      double eps, int timestep) {
    return guard([&] {
        check_sim(s);
        if (layer < 0 || layer >= s->L) throw Error(DMPM_ERR_INDEX, "Layer index out of range.");
        descend_adam(s, layer, alpha, grad_norm, beta1, beta2, eps, timestep);
    });
}

int dmpm_optimize_single_timestep(dmpm_sim* s, int t, int max_gd_iters, int /*episode*/, double initial_alpha,
                                  int max_ls) {
    return guard([&] {
        check_sim(s);
        s->refuse_if_captured("dmpm_optimize_single_timestep");
        if (t < 0 || t >= s->L - 1) {
            fprintf(stderr, "Invalid timestep index: %d\n", t);
            return;
        }
        const double beta1 = 0.9, beta2 = 0.999, eps = 1e-3;
        double alpha = initial_alpha;
        ensure_adam(s, t);
        void* bak = s->dalloc((size_t)4 * s->ctl_bytes());
        for (int gd = 0; gd < max_gd_iters; ++gd) {
            s->forward(t);
            const double gd_loss = end_layer_mass_loss(s);
            if (!std::isfinite(gd_loss)) {
                printf("Warning: Non-finite loss. Aborting.\n");
                break;
            }
            s->backward(t);
            const double gn = std::max(control_norm(s), 1e-12);
            backup_layer(s, t, bak);
            bool accepted = false;
            double alpha_try = alpha;
            for (int ls = 0; ls < max_ls; ++ls) {
                descend_adam(s, t, alpha_try, gn, beta1, beta2, eps, g_single_timestep_adam_t + 1);
                s->forward(t);
                const double nl = end_layer_mass_loss(s);
                if (std::isfinite(nl) && nl < gd_loss) {
                    g_single_timestep_adam_t++;
                    alpha = std::min(alpha_try * 1.1, initial_alpha);
                    accepted = true;
                    break;
                }
                restore_layer(s, t, bak);
                alpha_try *= 0.5;
            }
            if (!accepted) {
                printf("  Line search failed.\n");
                break;
            }
        }
        HIPCHK(hipStreamSynchronize(s->stream));
        s->dfree(bak);
    });
}

int dmpm_single_timestep_adam_count(int* count, int reset) {
    return guard([&] {
        if (count) *count = g_single_timestep_adam_t;
        if (reset) g_single_timestep_adam_t = 0;
    });
}

int dmpm_run_optimization(dmpm_sim* s, int num_steps, double dt, double drag, const double f_ext[3],
                          int control_stride, int max_gd_iters, int max_ls, double initial_alpha, double gd_tol,
                          double smoothing, int /*current_episodes*/, double* final_loss) {
    return guard([&] {
        check_sim(s);
        s->refuse_if_captured("dmpm_run_optimization");
        if (num_steps <= 0) throw Error(DMPM_ERR_ARG, "num_steps must be positive");
        if (control_stride <= 0) throw Error(DMPM_ERR_ARG, "control_stride must be positive");
        s->cfg.dt = dt;
        s->cfg.drag = drag;
        s->cfg.smoothing = smoothing;
        for (int i = 0; i < 3; ++i) s->cfg.f_ext[i] = f_ext ? f_ext[i] : 0.0;
        s->ph = phys_of(s->cfg);
        if (dmpm_set_up_comp_graph(s, num_steps) != DMPM_OK) throw Error(DMPM_ERR_STATE, g_err);
        s->forward(0);
        double loss = end_layer_mass_loss(s);
        s->backward(0);
        const double beta1 = 0.9, beta2 = 0.999, eps = 1e-3;
        int adam_t = 0;
        void* bak = s->dalloc((size_t)4 * s->ctl_bytes());
        const int passes = 3;  // CompGraph.cpp:246
        for (int pass = 0; pass < passes; ++pass) {
            for (int ct = 0; ct < num_steps - 1; ct += control_stride) {
                double alpha = initial_alpha;
                double norm0 = 0;
                ensure_adam(s, ct);
                for (int gd = 0; gd < max_gd_iters; ++gd) {
                    s->forward(ct);
                    const double gd_loss = end_layer_mass_loss(s);
                    if (!std::isfinite(gd_loss)) {
                        printf("Warning: Non-finite loss detected. Aborting step.\n");
                        break;
                    }
                    s->backward(ct);
                    const double gn = std::max(control_norm(s), 1e-12);
                    if (gd == 0) norm0 = gn;
                    if (gn < gd_tol * norm0) break;
                    backup_layer(s, ct, bak);
                    bool accepted = false;
                    double alpha_try = alpha;
                    for (int ls = 0; ls < max_ls; ++ls) {
                        descend_adam(s, ct, alpha_try, gn, beta1, beta2, eps, adam_t + 1);
                        s->forward(ct);
                        const double nl = end_layer_mass_loss(s);
                        if (std::isfinite(nl) && nl < gd_loss) {
                            adam_t++;
                            alpha = std::min(alpha_try * 1.1, initial_alpha);
                            accepted = true;
                            break;
                        }
                        restore_layer(s, ct, bak);
                        alpha_try *= 0.5;
                    }
                    if (!accepted) break;
                }
            }
        }
        loss = end_layer_mass_loss(s);
        if (final_loss) *final_loss = loss;
        s->dfree(bak);
    });
}

int dmpm_p2g_mass(const dmpm_config* cfg, int n, const void* x, const void* m, void* grid_m, int dtype,
                  int on_device) {
    return guard([&] { mass_volume_call(cfg, n, x, m, nullptr, grid_m, dtype, on_device); });
}

int dmpm_calculate_volumes(const dmpm_config* cfg, int n, const void* x, const void* m, void* vol, void* grid_m,
                           int dtype, int on_device) {
    return guard([&] {
        if (!vol) throw Error(DMPM_ERR_ARG, "null argument");
        mass_volume_call(cfg, n, x, m, vol, grid_m, dtype, on_device);
    });
}

int dmpm_lame_parameters(double E, double nu, double* lam, double* mu) {
    return guard([&] {
        if (!lam || !mu) throw Error(DMPM_ERR_ARG, "null output");
        const float e = (float)E, p = (float)nu;
        *lam = (double)(e * p / ((1.f + p) * (1.f - 2.f * p)));
        *mu = (double)(e / (2.f + 2.f * p));
    });
}

int dmpm_last_timing(const dmpm_sim* s, double* fwd_ms, double* bwd_ms) {
    return guard([&] {
        check_sim(s);
        float a = 0, b = 0;
        HIPCHK(hipEventSynchronize(s->ev[3]));
        if (hipEventElapsedTime(&a, s->ev[0], s->ev[1]) != hipSuccess) a = 0;
        if (hipEventElapsedTime(&b, s->ev[2], s->ev[3]) != hipSuccess) b = 0;
        (void)hipGetLastError();
        if (fwd_ms) *fwd_ms = a;
        if (bwd_ms) *bwd_ms = b;
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// hipGraph capture of whole episodes.  A dmpm_graph holds one instantiated
// graph with one branch per sim: each sim's forward pass (keeping controls)
// from layer `start` and its backward pass to layer `stop`, recorded on the
// sim's own stream, forked from and joined back to an origin stream.  The
// branches are independent, so the GPU runs the instances concurrently.
// Holds one count on each captured sim while alive.
// The sims' host bookkeeping after the captured episode is kept and put back
// at every launch; a launch is refused once one of the sims is destroyed.
struct CaptureHold {
    std::vector<std::shared_ptr<int>> counts;
    std::vector<std::shared_ptr<bool>> alive;
    std::vector<dmpm_sim*> sims;
    std::vector<dmpm_sim::Book> books;
    void hold(dmpm_sim** s, int n) {
        for (int i = 0; i < n; ++i) {
            counts.push_back(s[i]->captures);
            ++*counts.back();
            alive.push_back(s[i]->alive);
            sims.push_back(s[i]);
            books.push_back(s[i]->book());
        }
    }
    // before a launch
    void renew() {
        for (auto& a : alive)
            if (!*a) throw Error(DMPM_ERR_STATE, "a sim of this capture has been destroyed");
        for (size_t i = 0; i < sims.size(); ++i) sims[i]->restore(books[i]);
    }
    ~CaptureHold() {
        for (auto& c : counts) --*c;
    }
};

struct dmpm_graph {
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    int count = 0;
    CaptureHold hold;
    ~dmpm_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : events) (void)hipEventDestroy(e);
        if (origin) (void)hipStreamDestroy(origin);
    }
};

extern "C" {

int dmpm_graph_capture(dmpm_sim** sims, int count, int start_layer, int stop_layer, dmpm_graph** out) {
    return guard([&] {
        if (!sims || count <= 0 || !out) throw Error(DMPM_ERR_ARG, "bad arguments");
        auto g = std::make_unique<dmpm_graph>();
        g->count = count;
        HIPCHK(hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking));
        g->events.resize(count + 1);
        for (auto& e : g->events) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        for (int i = 0; i < count; ++i) {
            dmpm_sim* s = sims[i];
            check_sim(s);
            if (s->n <= 0) throw Error(DMPM_ERR_STATE, "sim without particles");
            if (s->cfg.deterministic) throw Error(DMPM_ERR_STATE, "deterministic mode cannot be captured");
            for (int j = 0; j < i; ++j)
                if (sims[j]->stream == s->stream) throw Error(DMPM_ERR_ARG, "sims must have distinct streams");
            s->ktimer = false;
            // everything the passes allocate or copy lazily must exist before capture
            for (int t = 0; t < s->L; ++t)
                if (s->want_gctl[t] && !s->gctl[t]) s->gctl[t] = s->dalloc(s->ctl_bytes());
            s->materialize_all();
            HIPCHK(hipStreamSynchronize(s->stream));
        }
        HIPCHK(hipStreamBeginCapture(g->origin, hipStreamCaptureModeRelaxed));
        hipError_t cap_err = hipSuccess;
        std::string msg;
        try {
            HIPCHK(hipEventRecord(g->events[count], g->origin));
            for (int i = 0; i < count; ++i) {
                dmpm_sim* s = sims[i];
                HIPCHK(hipStreamWaitEvent(s->stream, g->events[count], 0));
                s->capturing = true;
                s->forward(start_layer, false);
                s->backward(stop_layer);
                s->capturing = false;
                HIPCHK(hipEventRecord(g->events[i], s->stream));
                HIPCHK(hipStreamWaitEvent(g->origin, g->events[i], 0));
            }
        } catch (const std::exception& e) {
            msg = e.what();
            cap_err = hipErrorUnknown;
        }
        for (int i = 0; i < count; ++i) sims[i]->capturing = false;
        hipGraph_t graph = nullptr;
        hipError_t end_err = hipStreamEndCapture(g->origin, &graph);
        if (cap_err != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            throw Error(DMPM_ERR_HIP, "capture failed: " + msg);
        }
        HIPCHK(end_err);
        g->graph = graph;
        HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        g->hold.hold(sims, count);
        *out = g.release();
    });
}

int dmpm_graph_launch(dmpm_graph* g) {
    return guard([&] {
        if (!g || !g->exec) throw Error(DMPM_ERR_ARG, "null graph");
        g->hold.renew();
        HIPCHK(hipGraphLaunch(g->exec, g->origin));
    });
}

int dmpm_graph_synchronize(dmpm_graph* g) {
    return guard([&] {
        if (!g) throw Error(DMPM_ERR_ARG, "null graph");
        HIPCHK(hipStreamSynchronize(g->origin));
    });
}

int dmpm_graph_destroy(dmpm_graph* g) {
    return guard([&] { delete g; });
}

// ---- batched episodes ---------------------------------------------------------
struct dmpm_batch {
    int count = 0;
    CaptureHold hold;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* dev = nullptr;  // parameter blocks, block prefix sums and ranges of every op
    ~dmpm_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dev) (void)hipFree(dev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

int dmpm_batch_capture(dmpm_sim** sims, int count, int start_layer, int stop_layer, dmpm_batch** out) {
    return guard([&] {
        if (!sims || count <= 0 || !out) throw Error(DMPM_ERR_ARG, "bad arguments");
        for (int i = 0; i < count; ++i) {
            check_sim(sims[i]);
            if (sims[i]->n <= 0) throw Error(DMPM_ERR_STATE, "sim without particles");
            if (sims[i]->cfg.deterministic) throw Error(DMPM_ERR_STATE, "deterministic mode cannot be batched");
        }
        // a batched episode runs each sim in one storage order (its recorded
        // operations have no permutations between segments)
        for (int i = 0; i < count; ++i) {
            if (!sims[i]->oseg.empty()) sims[i]->refuse_if_captured("dmpm_batch_capture of a sim with segment orders");
            sims[i]->collapse_orders();
            sims[i]->materialize_all();
        }
        dmpm_sim* s0 = sims[0];
        const Launch L0 = s0->launch();
        for (int i = 1; i < count; ++i) {
            const dmpm_sim* s = sims[i];
            const Launch Li = s->launch();
            if (s->pr != s0->pr || s->cfg.order != s0->cfg.order || s->cfg.model != s0->cfg.model ||
                s->cfg.adjoint != s0->cfg.adjoint || Li.block != L0.block || Li.lds_bytes != L0.lds_bytes ||
                s->cfg.unfused != s0->cfg.unfused)
                throw Error(DMPM_ERR_ARG, "batched sims must share precision, B-spline order, model, adjoint mode, "
                                          "block size, LDS budget and fusion");
        }
        // record each sim's episode (host bookkeeping runs as for an eager episode)
        std::vector<std::vector<dmpm_sim::Op>> ops(count);
        for (int i = 0; i < count; ++i) {
            dmpm_sim* s = sims[i];
            HIPCHK(hipStreamSynchronize(s->stream));
            s->rec = &ops[i];
            try {
                s->forward(start_layer, false);
                s->backward(stop_layer);
            } catch (...) {
                s->rec = nullptr;
                throw;
            }
            s->rec = nullptr;
        }
        const size_t nops = ops[0].size();
        for (int i = 1; i < count; ++i) {
            if (ops[i].size() != nops) throw Error(DMPM_ERR_ARG, "batched sims must run episodes of the same structure");
            for (size_t k = 0; k < nops; ++k)
                if (ops[i][k].kind != ops[0][k].kind)
                    throw Error(DMPM_ERR_ARG, "batched sims must run episodes of the same structure");
        }
        // lay out every op's table in one buffer
        const Prec pr = s0->pr;
        // the kernels index the parameter table as an array of the kernels'
        // parameter struct, so the host stride must be exactly its size;
        // only each table's start is aligned
        const size_t kpb = kp_bytes(pr);
        const size_t esz = s0->esz;
        struct Item {
            size_t off, off2;
            int total;
            long long max_n;
        };
        std::vector<Item> items(nops);
        size_t bytes = 0;
        for (size_t k = 0; k < nops; ++k) {
            items[k].off = bytes;
            if (op_is_kernel(ops[0][k].kind)) {
                bytes += kpb * kp_count(ops[0][k].kind) * count;
                bytes = (bytes + 15) & ~size_t(15);
                items[k].off2 = bytes;
                bytes += ((count + 1) * sizeof(int) + 15) & ~size_t(15);
            } else {
                bytes += sizeof(Seg) * count;
                bytes = (bytes + 15) & ~size_t(15);
            }
        }
        std::vector<char> host(std::max<size_t>(bytes, 16), 0);
        for (size_t k = 0; k < nops; ++k) {
            const int kind = ops[0][k].kind;
            if (op_is_kernel(kind)) {
                int* bs = (int*)(host.data() + items[k].off2);
                bs[0] = 0;
                for (int i = 0; i < count; ++i) {
                    const Launch& Li = ops[i][k].L;
                    build_kp(pr, kind, Li, &ops[i][k].L2, host.data() + items[k].off + kpb * kp_count(kind) * i);
                    bs[i + 1] = bs[i] + instance_blocks(Li);
                }
                items[k].total = bs[count];
            } else {
                Seg* sg = (Seg*)(host.data() + items[k].off);
                long long mx = 0;
                for (int i = 0; i < count; ++i) {
                    const auto& o = ops[i][k];
                    if (kind != OP_ACCUM && ((((uintptr_t)o.dst | (uintptr_t)o.src) & 15) || (o.n * (long long)esz) % 16))
                        throw Error(DMPM_ERR_STATE, "unaligned buffer range in a batched episode");
                    sg[i] = Seg{o.dst, o.src, o.n};
                    mx = std::max(mx, o.n);
                }
                items[k].max_n = mx;
            }
        }
        auto b = std::make_unique<dmpm_batch>();
        b->count = count;
        HIPCHK(hipMalloc(&b->dev, host.size()));
        HIPCHK(hipMemcpy(b->dev, host.data(), host.size(), hipMemcpyHostToDevice));
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        hipError_t launch_err = hipSuccess;
        for (size_t k = 0; k < nops; ++k) {
            const int kind = ops[0][k].kind;
            const char* base = (const char*)b->dev;
            if (op_is_kernel(kind))
                launch_batch(pr, kind, s0->ph, L0.lds_bytes, L0.block, base + items[k].off,
                             (const int*)(base + items[k].off2), count, items[k].total, b->stream);
            else
                launch_segs(pr, kind, (const Seg*)(base + items[k].off), count, items[k].max_n, b->stream);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess && launch_err == hipSuccess) launch_err = e;
        }
        hipGraph_t graph = nullptr;
        const hipError_t end_err = hipStreamEndCapture(b->stream, &graph);
        b->graph = graph;
        HIPCHK(launch_err);
        HIPCHK(end_err);
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->hold.hold(sims, count);
        *out = b.release();
    });
}

int dmpm_batch_launch(dmpm_batch* b) {
    return guard([&] {
        if (!b || !b->exec) throw Error(DMPM_ERR_ARG, "null batch");
        b->hold.renew();
        HIPCHK(hipGraphLaunch(b->exec, b->stream));
    });
}

int dmpm_batch_synchronize(dmpm_batch* b) {
    return guard([&] {
        if (!b) throw Error(DMPM_ERR_ARG, "null batch");
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int dmpm_batch_destroy(dmpm_batch* b) {
    return guard([&] { delete b; });
}

int dmpm_kernel_timer(dmpm_sim* s, int capacity) {
    return guard([&] {
        check_sim(s);
        HIPCHK(hipStreamSynchronize(s->stream));
        s->krec.clear();
        s->knext = 0;
        s->koverflow = 0;
        s->ktimer = capacity > 0;
        const size_t want = 2 * (size_t)std::max(capacity, 0);
        while (s->kpool.size() < want) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            s->kpool.push_back(e);
        }
    });
}

int dmpm_kernel_stats(dmpm_sim* s, int kind, double* total_ms, int* count) {
    return guard([&] {
        check_sim(s);
        if (kind < 0 || kind >= dmpm_sim::NKIND) throw Error(DMPM_ERR_ARG, "unknown kernel kind");
        HIPCHK(hipStreamSynchronize(s->stream));
        double tot = 0;
        int c = 0;
        for (auto& r : s->krec) {
            if (r.first != kind) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, s->kpool[r.second], s->kpool[r.second + 1]));
            tot += ms;
            ++c;
        }
        if (total_ms) *total_ms = tot;
        if (count) *count = c;
    });
}

int dmpm_window_counters(dmpm_sim* s, int on) {
    return guard([&] {
        check_sim(s);
        s->count_windows = on != 0;
    });
}

int dmpm_window_stats(dmpm_sim* s, long long* lds_blocks, long long* fallback_blocks, int reset) {
    return guard([&] {
        check_sim(s);
        unsigned long long h[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(h, s->wstat, sizeof(h), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        if (lds_blocks) *lds_blocks = (long long)h[0];
        if (fallback_blocks) *fallback_blocks = (long long)h[1];
        if (reset) s->zero(s->wstat, 4);
    });
}

int dmpm_memory(const dmpm_sim* s, size_t* current, size_t* peak) {
    return guard([&] {
        check_sim(s);
        if (current) *current = s->mem_cur;
        if (peak) *peak = s->mem_peak;
    });
}

}  // extern "C"
