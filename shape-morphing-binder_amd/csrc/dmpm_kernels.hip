// MPM step kernels for gfx950 (MI355X).
//
// One thread per particle, 256-thread blocks (4 wave64s).  Particles are kept
// in Hilbert order of their stencil base cell, so a block's particles occupy
// a compact box of the grid.  Every particle<->grid kernel first reduces that
// box (the "window": the bounding box of the block's B-spline stencils) with
// wave shuffles, then
//   * scatters (P2G, adjoint of G2P) sum each node value over runs of lanes
//     that share a base cell, accumulate the run totals into the window in
//     LDS (FP64) and flush it to HBM with one global atomic per touched node
//     value, instead of one per particle-node pair;
//   * gathers (G2P, adjoint of P2G) stage the window's node values in LDS
//     once and read the stencil from there.
// When a block's window does not fit the LDS budget (scattered particles),
// the block falls back to direct global atomics / global gathers, so the
// result never depends on the particle order, only the speed does.
//
// Grid velocity is never stored: v_i = p_i/m_i + dt f_ext (0 when m_i <= 1e-12
// or on a wall node) is recomputed from (m, p) wherever it is read, which is
// the reference's grid update (ForwardSimulation.cpp:82-87, 165-170) fused
// into its readers.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <stdexcept>

#include <hipcub/hipcub.hpp>

#include "dmpm_kernels.h"
#include "dmpm_math.h"

// The build compiles this file once per part, in parallel, with DMPM_PART
// set: part 0 holds the small kernels and every host entry point declared in
// dmpm_kernels.h; parts 1-6 each instantiate one group of the big particle
// kernels for one precision (1/2: forward FP32/FP64, 3/4: adjoint scatter and
// gather, 5/6: fused adjoint), reached from part 0 through the launch_*<T>
// templates declared below.  Without DMPM_PART one unit holds everything.
#ifndef DMPM_PART
#define DMPM_PART -1
#endif
#define DMPM_IN(p) (DMPM_PART < 0 || DMPM_PART == (p))

namespace dmpm {

// Per-part launchers (defined and instantiated in parts 1-6).
template <typename T> void launch_fwd(int kind, const Launch& L, hipStream_t s);
template <typename T> void launch_adj(int kind, const Launch& L, hipStream_t s);
template <typename T> void launch_fused(const Launch& L, const Launch& L2, hipStream_t s);
// One batched launch of kernel op `kind` over `count` instances (parts 1-6 by
// kind group and precision).
template <typename T>
void batch_fwd(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab, const int* bstart,
               int count, hipStream_t s);
template <typename T>
void batch_adj(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab, const int* bstart,
               int count, hipStream_t s);
template <typename T>
void batch_fused(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab, const int* bstart,
                 int count, hipStream_t s);

template <typename T>
struct KP {
    int n;
    long long ld, gld;
    int nx, ny, nz;
    T dx, mn0, mn1, mn2;
    T dt, drag, sm, zeta, jmin, f0, f1, f2;
    int trust, walls;
    const T* st;
    const T* stat;
    const T* ctl;
    T* out;
    T* gm;   // grid: 4 values per node (m, p0, p1, p2); a mass-only P2G's grid: 1 (m)
    T* ggv;  // adjoint grid dL/dv_i: 4 values per node (the last unused)
    const T* adj1;
    T* adj0;
    T* gctl;
    const T* ctl1;  // control of layer t+1 (fused step)
    T* gm1;         // grid t+1 (fused step), laid out as gm
    const T* nxt;   // adjoint: layer t+1 state (v, C read instead of a G2P) or null
    const int* opos;  // output row of each particle in another storage order, or null
    const int* bst; // particle block table (nblk + 1 starts) or null
    int nblk;
    unsigned long long* wstat;
    uint4* zero;       // cleared by this launch for a later one (or null)
    long long zero16;  // its 16-byte chunks
    long long zper;    // chunks per logical block: ceil(zero16 / zblk)
    int zblk;          // logical blocks sharing the clear
    int cap;
    int park;          // adjoint gather: byte offset in LDS of the values parked through its node loop
};

static inline unsigned grid_blocks(int n, int block) {
    unsigned nb = (unsigned)((n + block - 1) / block);
    if (nb == 0) nb = 1;
    return (nb + 7u) & ~7u;
}

// Launch grid of a window kernel: the block table's blocks, or n / block.
static inline unsigned launch_blocks(const Launch& L) {
    if (!L.bst) return grid_blocks(L.n, L.block);
    const unsigned nb = (unsigned)std::max(L.nblk, 1);
    return (nb + 7u) & ~7u;
}

template <typename T>
static KP<T> make_kp(const Launch& L, int narrays) {
    KP<T> k;
    k.n = L.n;
    k.ld = L.ld;
    k.gld = L.gld;
    k.nx = L.ph.dims[0];
    k.ny = L.ph.dims[1];
    k.nz = L.ph.dims[2];
    k.dx = (T)L.ph.dx;
    k.mn0 = (T)L.ph.grid_min[0];
    k.mn1 = (T)L.ph.grid_min[1];
    k.mn2 = (T)L.ph.grid_min[2];
    k.dt = (T)L.ph.dt;
    k.drag = (T)L.ph.drag;
    k.sm = (T)L.ph.smoothing;
    k.zeta = (T)L.ph.zeta;
    k.jmin = (T)L.ph.jmin;
    k.f0 = (T)L.ph.f_ext[0];
    k.f1 = (T)L.ph.f_ext[1];
    k.f2 = (T)L.ph.f_ext[2];
    k.trust = L.ph.trust_region;
    k.walls = L.ph.wall_cells;
    k.st = (const T*)L.st;
    k.stat = (const T*)L.stat;
    k.ctl = (const T*)L.ctl;
    k.out = (T*)L.out;
    k.gm = (T*)L.grid;
    k.ggv = (T*)L.ggv;
    k.adj1 = (const T*)L.adj1;
    k.adj0 = (T*)L.adj0;
    k.gctl = (T*)L.gctl;
    k.ctl1 = (const T*)L.ctl1;
    k.gm1 = (T*)L.grid1;
    k.nxt = (const T*)L.nxt;
    k.opos = L.opos;
    k.bst = L.bst;
    k.nblk = L.nblk;
    k.wstat = L.wstat;
    k.zero = (uint4*)L.zero;
    k.zero16 = L.zero ? L.zero_bytes / 16 : 0;
    k.zblk = (int)launch_blocks(L);
    k.zper = (k.zero16 + k.zblk - 1) / k.zblk;
    k.cap = narrays > 0 ? (int)(L.lds_bytes / narrays) : 0;  // narrays = LDS bytes per window node
    k.park = 0;
    return k;
}

// Blocks are handed out round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remap so that each XCD gets a contiguous run of logical blocks:
// spatially adjacent particle blocks then share that XCD's L2, which holds
// the halo nodes their windows overlap on.  The launch grid is padded to a
// multiple of 8; logical blocks past the end have no particles.
DMPM_DI int logical_block() {
    const int nb = gridDim.x;
    const int per = nb >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Blocks of a batched launch: instance `inst` owns the logical blocks
// [bstart[inst], bstart[inst + 1]) (logical order as in logical_block(); the
// grid is padded to a multiple of 8 and blocks past bstart[count] have no
// work).  Block-uniform, so the search runs on the scalar unit.
DMPM_DI bool batch_block(const int* bstart, int count, int& inst, int& lblock) {
    const int lb = logical_block();
    if (lb >= bstart[count]) return false;
    int lo = 0, hi = count;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (bstart[mid] <= lb) lo = mid;
        else hi = mid;
    }
    inst = lo;
    lblock = lb - bstart[lo];
    return true;
}

// A launch's blocks clear, after their own work, a buffer that a later
// launch on the same stream accumulates into, so that the clear takes no
// launch of its own and its stores overlap the launch's arithmetic: logical
// block `lblock` of a.zblk clears its share of the 16-byte chunks, with
// non-temporal stores (the lines are not read again by this launch; measured
// slightly faster than plain stores, and than clearing before the block's
// own work).
template <typename T>
DMPM_DI void clear_share(const KP<T>& a, int lblock) {
    if (!a.zero) return;
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const long long per = a.zper;  // (a 64-bit division on the scalar unit per wave otherwise)
    const long long c1 = min(a.zero16, (long long)(lblock + 1) * per);
    u32x4* z = reinterpret_cast<u32x4*>(a.zero);
    for (long long c = (long long)lblock * per + threadIdx.x; c < c1; c += blockDim.x)
        __builtin_nontemporal_store((u32x4)(0u), z + c);
}

// Particles [i0, i1) of logical block `lblock` (block-uniform).  Logical
// blocks past the end (the grid is padded to a multiple of 8) are empty.
template <typename T>
DMPM_DI void block_range(const KP<T>& a, int lblock, int& i0, int& i1) {
    if (a.bst) {
        if (lblock < a.nblk) {
            i0 = a.bst[lblock];
            i1 = a.bst[lblock + 1];
        } else {
            i0 = i1 = a.n;
        }
    } else {
        i0 = lblock * (int)blockDim.x;
        i1 = min(a.n, i0 + (int)blockDim.x);
    }
}

// A block's window: the box [lo, lo + e) of grid nodes its particles'
// stencils touch, laid out in LDS with node (i, j, k) of the box at
// i*s1 + j*s2 + k.  The strides are padded to odd numbers of nodes, so the
// same stencil offset of particles in neighbouring cells (what one LDS
// instruction of a wave addresses) spreads over the banks instead of
// landing on one every few rows; vol counts the padded slots.
struct Win {
    int lo0, lo1, lo2, e1, e2, s1, s2, vol;
    unsigned r1, r2;  // ceil(2^32 / s1), ceil(2^32 / s2): slot -> node without a division
    bool lds;
    bool any;
};

// Wave-wide min of a0, a1, a2 and max of b0, b1, b2, left in lane 63,
// through DPP: quad swaps, half-row and row mirrors, then row broadcasts 15
// and 31 fold the four 16-lane rows (rows left out by the row mask keep
// their value).  The min / max takes its first operand through the DPP
// modifier itself, one instruction per step and value, and the six
// reductions are interleaved so that no DPP operand was written by one of
// the five instructions before it.  All lanes must be active.
#define DMPM_BOX_STEP(CTRL)                            \
    "v_min_i32_dpp %0, %0, %0 " CTRL "\n"             \
    "v_min_i32_dpp %1, %1, %1 " CTRL "\n"             \
    "v_min_i32_dpp %2, %2, %2 " CTRL "\n"             \
    "v_max_i32_dpp %3, %3, %3 " CTRL "\n"             \
    "v_max_i32_dpp %4, %4, %4 " CTRL "\n"             \
    "v_max_i32_dpp %5, %5, %5 " CTRL "\n"
DMPM_DI void wave_box(int& a0, int& a1, int& a2, int& b0, int& b1, int& b2) {
    asm volatile(
        "s_nop 4\n"
        DMPM_BOX_STEP("quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf")
        DMPM_BOX_STEP("quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf")
        DMPM_BOX_STEP("row_half_mirror row_mask:0xf bank_mask:0xf")
        DMPM_BOX_STEP("row_mirror row_mask:0xf bank_mask:0xf")
        DMPM_BOX_STEP("row_bcast:15 row_mask:0xa bank_mask:0xf")
        DMPM_BOX_STEP("row_bcast:31 row_mask:0xc bank_mask:0xf")
        "s_nop 4"
        : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(b0), "+v"(b1), "+v"(b2));
}
#undef DMPM_BOX_STEP

// Bounding box of the block's stencils (base nodes b, width KS).  Every
// thread of the block must call this (all lanes of each wave active).
template <int KS>
DMPM_DI Win block_window(bool valid, int b0, int b1, int b2, int cap) {
    __shared__ int s_red[6][16];
    int v0 = valid ? b0 : INT_MAX, v1 = valid ? b1 : INT_MAX, v2 = valid ? b2 : INT_MAX;
    int w0 = valid ? b0 : INT_MIN, w1 = valid ? b1 : INT_MIN, w2 = valid ? b2 : INT_MIN;
    wave_box(v0, v1, v2, w0, w1, w2);
    v0 = __builtin_amdgcn_readlane(v0, 63);
    v1 = __builtin_amdgcn_readlane(v1, 63);
    v2 = __builtin_amdgcn_readlane(v2, 63);
    w0 = __builtin_amdgcn_readlane(w0, 63);
    w1 = __builtin_amdgcn_readlane(w1, 63);
    w2 = __builtin_amdgcn_readlane(w2, 63);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s_red[0][wave] = v0; s_red[1][wave] = v1; s_red[2][wave] = v2;
        s_red[3][wave] = w0; s_red[4][wave] = w1; s_red[5][wave] = w2;
    }
    __syncthreads();
    const int nw = blockDim.x >> 6;
    int l0 = s_red[0][0], l1 = s_red[1][0], l2 = s_red[2][0];
    int h0 = s_red[3][0], h1 = s_red[4][0], h2 = s_red[5][0];
    for (int q = 1; q < nw; ++q) {
        l0 = min(l0, s_red[0][q]); l1 = min(l1, s_red[1][q]); l2 = min(l2, s_red[2][q]);
        h0 = max(h0, s_red[3][q]); h1 = max(h1, s_red[4][q]); h2 = max(h2, s_red[5][q]);
    }
    Win w;
    w.any = l0 != INT_MAX;
    w.lo0 = l0; w.lo1 = l1; w.lo2 = l2;
    if (!w.any) {
        w.e1 = w.e2 = w.s1 = w.s2 = 0; w.vol = 0; w.lds = true;
        w.r1 = w.r2 = 0u;
        return w;
    }
    const long long e0 = (long long)h0 - l0 + KS, e1 = (long long)h1 - l1 + KS, e2 = (long long)h2 - l2 + KS;
    // (extents are checked one by one first: base cells of far-away or
    // non-finite positions must not overflow the product)
    const long long s2 = e2 | 1, s1 = (e1 * s2) | 1;
    const long long vol = e0 <= cap && e1 <= cap && e2 <= cap ? e0 * s1 : LLONG_MAX;
    w.lds = vol <= (long long)cap;
    w.e1 = (int)e1;
    w.e2 = (int)e2;
    w.s1 = w.lds ? (int)s1 : 0;
    w.s2 = w.lds ? (int)s2 : 0;
    w.vol = w.lds ? (int)vol : 0;
    // 0xffffffff / s + 1 = ceil(2^32 / s) for s >= 2 (s2 >= KS >= 3 here)
    w.r1 = w.lds ? 0xffffffffu / (unsigned)w.s1 + 1u : 0u;
    w.r2 = w.lds ? 0xffffffffu / (unsigned)w.s2 + 1u : 0u;
    return w;
}

// Grid node of window slot t; false for the padding slots of the strides.
// The quotients come from the multipliers r = ceil(2^32 / s): with t = q s + e
// (0 <= e < s), t r / 2^32 = q + e / s + t (r - 2^32 / s) / 2^32, and the last
// two terms stay below 1 while t s < 2^32 (t and s are below 2^16: windows
// hold at most 160 KiB / 8 B nodes and callers pass t < vol + 4 * blockDim.x).
DMPM_DI bool window_node(const Win& w, int t, int& gi, int& gj, int& gk) {
    const int i = (int)__umulhi((unsigned)t, w.r1), r = t - i * w.s1;
    const int j = (int)__umulhi((unsigned)r, w.r2), k = r - j * w.s2;
    gi = w.lo0 + i;
    gj = w.lo1 + j;
    gk = w.lo2 + k;
    return j < w.e1 && k < w.e2;
}

// Count a scatter window as held in LDS or fallen back to global atomics.
template <typename T>
DMPM_DI void count_window(const KP<T>& a, const Win& w) {
    if (a.wstat && threadIdx.x == 0 && w.any) atomicAdd(a.wstat + (w.lds ? 0 : 1), 1ull);
}

template <typename T>
DMPM_DI bool is_wall(const KP<T>& a, int i, int j, int k) {
    const int w = a.walls;
    return w > 0 && (i < w || j < w || k < w || i >= a.nx - w || j >= a.ny - w || k >= a.nz - w);
}

template <typename T>
DMPM_DI bool in_grid(const KP<T>& a, int i, int j, int k) {
    return (unsigned)i < (unsigned)a.nx && (unsigned)j < (unsigned)a.ny && (unsigned)k < (unsigned)a.nz;
}

template <typename T>
DMPM_DI long long gindex(const KP<T>& a, int i, int j, int k) {
    return ((long long)i * a.ny + j) * a.nz + k;
}

// Particle planes are read and written with raw buffer instructions: one
// buffer resource per array (SGPRs), the particle's byte offset i * sizeof(T)
// in one VGPR shared by every plane, and the plane's byte offset
// plane * ld * sizeof(T) in an SGPR.  A 64-bit vector address per plane
// access (the global form) costs VALU instructions and two VGPRs each.
// Offsets are 32-bit: an array holds fewer than 2^32 bytes (the host checks).
DMPM_DI __amdgpu_buffer_rsrc_t plane_rsrc(const void* base) {
    // stride 0, num_records 0xffffffff, dword 3: 32-bit data format (gfx9)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
template <typename T>
DMPM_DI T load_s(const T* base, long long ld, int plane, int i) {
    const __amdgpu_buffer_rsrc_t r = plane_rsrc(base);
    const int vo = i * (int)sizeof(T), so = (int)(plane * ld * (long long)sizeof(T));
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
    }
}
template <typename T>
DMPM_DI void store_s(T* base, long long ld, int plane, int i, T v) {
    const __amdgpu_buffer_rsrc_t r = plane_rsrc(base);
    const int vo = i * (int)sizeof(T), so = (int)(plane * ld * (long long)sizeof(T));
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, vo, so, 0);
    } else {
        typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, vo, so, 0);
    }
}

// Grid arrays hold kNode values per node, node g at [kNode g, kNode g +
// kNode): m, p0, p1, p2 in the forward grid, dL/dv_i 0..2 and an unused slot
// in the adjoint grid.  Device-scope float atomics travel past the L2 to
// memory in 64-byte requests, so the window flushes, whose lanes walk a
// window row node by node and value by value, cover whole lines with one
// request; and a node is one 16-byte load.  A mass-only P2G writes one value per node.
constexpr int kNode = 4;

// Byte offset i * size + add of an array below 4 GiB (the host checks), in
// unsigned arithmetic: the buffer instructions take it as an unsigned 32-bit
// value, and it may exceed INT_MAX.
DMPM_DI int byte_offset(int i, unsigned size, unsigned add) {
    return (int)((unsigned)i * size + add);
}

// The node's values as kNode scalar buffer loads, which the compiler merges
// into one 16-byte (FP32) or two (FP64) loads.  (Element extraction from the
// vector the 64- and 128-bit buffer-load builtins return is narrowed to a
// single dword load by this compiler's optimiser, and is not used.)
template <typename T>
DMPM_DI void load_node(const T* base, int g, T* v) {
    const __amdgpu_buffer_rsrc_t r = plane_rsrc(base);
#pragma unroll
    for (int q = 0; q < kNode; ++q) {
        if constexpr (sizeof(T) == 4) {
            v[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_offset(g, 16, 4 * q), 0, 0));
        } else {
            v[q] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_offset(g, 32, 8 * q), 0, 0));
        }
    }
}

// Grid element e += v, a device-scope float atomic (buffer form for FP32;
// FP64 keeps the global form, which has no buffer counterpart).
template <typename T>
DMPM_DI void atomic_add_at(T* base, int e, T v) {
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(v, plane_rsrc(base), byte_offset(e, 4, 0), 0, 0);
    } else {
        unsafeAtomicAdd(base + e, v);
    }
}

// Grid velocity update (ForwardSimulation.cpp:82-87) of a node whose mass
// and momentum are loaded; `live` is false for walls and nodes off the grid.
// Selects rather than branches, so that the loads stay ahead of any use.
template <typename T>
DMPM_DI void velocity_of(const KP<T>& a, bool live, T m, T p0, T p1, T p2, T& v0, T& v1, T& v2) {
    const T inv = T(1) / m;
    live = live && m > T(1e-12);
    v0 = live ? p0 * inv + a.dt * a.f0 : T(0);
    v1 = live ? p1 * inv + a.dt * a.f1 : T(0);
    v2 = live ? p2 * inv + a.dt * a.f2 : T(0);
}

// Grid velocity update (ForwardSimulation.cpp:82-87) plus sticky walls.
template <typename T>
DMPM_DI void node_velocity(const KP<T>& a, long long g, int i, int j, int k, T& v0, T& v1, T& v2) {
    T n[kNode];
    load_node(a.gm, (int)g, n);
    velocity_of(a, !is_wall(a, i, j, k), n[0], n[1], n[2], n[3], v0, v1, v2);
}

template <typename T>
DMPM_DI m3<T> load_m3(const T* base, long long ld, int plane, int i) {
    m3<T> M;
#pragma unroll
    for (int q = 0; q < 9; ++q) M.a[q] = load_s(base, ld, plane + q, i);
    return M;
}
template <typename T>
DMPM_DI v3<T> load_v3(const T* base, long long ld, int plane, int i) {
    v3<T> v;
#pragma unroll
    for (int q = 0; q < 3; ++q) v.a[q] = load_s(base, ld, plane + q, i);
    return v;
}
template <typename T>
DMPM_DI void store_m3(T* base, long long ld, int plane, int i, const m3<T>& M) {
#pragma unroll
    for (int q = 0; q < 9; ++q) store_s(base, ld, plane + q, i, M.a[q]);
}
template <typename T>
DMPM_DI void store_v3(T* base, long long ld, int plane, int i, const v3<T>& v) {
#pragma unroll
    for (int q = 0; q < 3; ++q) store_s(base, ld, plane + q, i, v.a[q]);
}

// Per-particle stencil: base node and per-axis weights.
template <typename T, int ORDER>
struct Sten {
    static constexpr int KS = Spline<ORDER>::K;
    int b[3];
    T u[3][KS], w[3][KS], dw[3][KS];
    DMPM_DI void init(const KP<T>& a, const v3<T>& x) {
        const T inv = T(1) / a.dx;
        b[0] = Spline<ORDER>::weights((x.a[0] - a.mn0) * inv, u[0], w[0], dw[0]);
        b[1] = Spline<ORDER>::weights((x.a[1] - a.mn1) * inv, u[1], w[1], dw[1]);
        b[2] = Spline<ORDER>::weights((x.a[2] - a.mn2) * inv, u[2], w[2], dw[2]);
    }
};

// Trust region of P_op_2 (ForwardSimulation.cpp:136-152).
template <typename T>
struct Trust {
    T s;        // scale applied to C
    int state;  // 0 inactive, 1 scaled, 2 clamped to zero
    T denom, J0;
};

// No trust-region scaling (C' used as it is).
template <typename T>
DMPM_DI Trust<T> trust_none() {
    Trust<T> t;
    t.s = T(1);
    t.state = 0;
    t.denom = T(0);
    t.J0 = T(1);
    return t;
}

// The adjoint of step t needs G2P's v' and C' (before the trust region).
// Layer t+1 (a.nxt) stores v' and the C' after the trust region; that is all
// the reference's adjoint reads (BackPropagation.cpp:27-31 use layer n+1's C),
// and all the exact adjoint reads when the trust region is off.  Otherwise
// (or without a.nxt) they are recomputed by a G2P from grid t.
template <typename T, bool REF>
DMPM_DI bool reads_next(const KP<T>& a) {
    return a.nxt != nullptr && (REF || !a.trust);
}

template <typename T>
DMPM_DI Trust<T> trust_region(const KP<T>& a, const m3<T>& Ft, const m3<T>& Craw) {
    Trust<T> t;
    t.s = T(1);
    t.state = 0;
    t.denom = T(0);
    t.J0 = (T)fmax(mdet(mcast<double>(Ft)), 1e-6);
    if (a.trust) {
        const T trC = mtrace(Craw);
        if (trC < T(0)) {
            const T denom = fmin(T(-1e-6), a.dt * trC);
            T sv = (a.jmin / t.J0 - T(1)) / denom;
            if (sv < T(1)) {
                t.state = sv < T(0) ? 2 : 1;
                t.s = fmin(fmax(sv, T(0)), T(1));
                t.denom = denom;
            }
        }
    }
    return t;
}

// ---------------------------------------------------------------------------
// P2G pieces (ForwardSimulation.cpp:26-80) with the stress of P_op_1 (:16-24).

// Affine momentum matrix A_p = -D^{-1} dt vol (P Ft^T + zeta tr(C) I) + m C
// and the drag-scaled momentum m (1 - dt drag) v of one particle.
template <typename T, int ORDER, int MODEL>
DMPM_DI void p2g_terms(const KP<T>& a, T vol, T lam, T mu, const v3<T>& v, const m3<T>& Ft, const m3<T>& C, T m,
                       m3<T>& A, v3<T>& mv) {
    m3<T> str = Model<MODEL>::kirchhoff(Ft, lam, mu);  // P Ft^T
    const T tau = a.zeta * mtrace(C);
    str.a[0] += tau; str.a[4] += tau; str.a[8] += tau;
    const T k0 = -Spline<ORDER>::dinv(a.dx) * a.dt * vol;
#pragma unroll
    for (int q = 0; q < 9; ++q) A.a[q] = k0 * str.a[q] + m * C.a[q];
    const T dragf = T(1) - a.dt * a.drag;
    mv = {{m * v.a[0] * dragf, m * v.a[1] * dragf, m * v.a[2] * dragf}};
}

template <typename T, int ORDER, int MODEL>
DMPM_DI void p2g_terms(const KP<T>& a, int i, const v3<T>& v, const m3<T>& Ft, const m3<T>& C, T m, m3<T>& A,
                       v3<T>& mv) {
    p2g_terms<T, ORDER, MODEL>(a, load_s(a.stat, a.ld, SVOL, i), load_s(a.stat, a.ld, SLAM, i), load_s(a.stat, a.ld, SMU, i), v, Ft,
                               C, m, A, mv);
}

// Type of the LDS scatter accumulators.  ds_add_f64 sustains about twice
// the throughput of ds_add_f32 on gfx950 under the same access pattern
// (measured), so windows accumulate in FP64 for both precisions; the
// flush converts once per node.
using Acc = double;

// Select a[i] from a small register array without dynamic indexing.

// ---------------------------------------------------------------------------
// Segmented reduction across lanes that share a base cell.  Particles are in
// Hilbert order, so lanes of a wave with the same base cell -- and therefore
// the same stencil nodes -- are adjacent.  Their contributions to each node
// are summed across lanes with DPP row shifts (inside 16-lane rows, 4
// Hillis-Steele steps) and only the head lane of each run issues the LDS
// atomic: the atomic count drops by the run length (~4 at 4-5 particles per
// cell), and LDS atomics are what bounds the scatter.  (Runs inside 16-lane
// rows rather than 8-lane chunks cost one more step and measured 7 % faster
// on the fused step.)

// lane i receives v of lane i+D (D = 1, 2, 4, 8); lanes whose source lies past
// their 16-lane row receive 0, which is never used: such a lane's run ends
// before the source.
template <int D>
DMPM_DI float dpp_down(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x100 + D, 0xF, 0xF, true));
}
template <int D>
DMPM_DI double dpp_down(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x100 + D, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + D, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
DMPM_DI int dpp_next(int k) { return __builtin_amdgcn_update_dpp(0, k, 0x101, 0xF, 0xF, true); }

// Which Hillis-Steele steps a lane takes (as 0/1 factors) and whether it
// heads its run.
template <typename T>
struct SegRun {
    T m1, m2, m4, m8;
    bool head;
};

// Runs are maximal sets of adjacent valid lanes inside one 16-lane row with
// the same base cell b.  Every lane of the wave must call this, and every
// cross-lane read below runs unconditionally: a read from a lane that is
// switched off returns 0, which would split a run on one side only.  A lane
// heads its run exactly when its left neighbour ends one, so the two views
// always agree.
template <typename T>
DMPM_DI SegRun<T> seg_run(bool valid, const int* b) {
    const int lane = threadIdx.x & 63;
    const int vf = valid ? 1 : 0;
    const int nv = dpp_next(vf);
    const int n0 = dpp_next(b[0]);
    const int n1 = dpp_next(b[1]);
    const int n2 = dpp_next(b[2]);
    const bool same_next = valid & (nv != 0) & (n0 == b[0]) & (n1 == b[1]) & (n2 == b[2]);
    // f: the lane's run ends within [lane, lane + d - 1]
    int f = (((lane & 15) == 15) | !same_next) ? 1 : 0;
    const int end_prev = __builtin_amdgcn_update_dpp(1, f, 0x111, 0xF, 0xF, false);
    SegRun<T> r;
    r.head = ((lane & 15) == 0) | (end_prev != 0);
    r.m1 = f ? T(0) : T(1);
    f |= __builtin_amdgcn_update_dpp(1, f, 0x101, 0xF, 0xF, false);
    r.m2 = f ? T(0) : T(1);
    f |= __builtin_amdgcn_update_dpp(1, f, 0x102, 0xF, 0xF, false);
    r.m4 = f ? T(0) : T(1);
    f |= __builtin_amdgcn_update_dpp(1, f, 0x104, 0xF, 0xF, false);
    r.m8 = f ? T(0) : T(1);
    return r;
}

// Sum v over the lane's run from this lane to the run's end; the head lane
// ends with the run total.  A step a lane does not take adds 0 * (the value
// read), so every lane's v must be finite.  Every lane of the wave must call
// this.
template <typename T>
DMPM_DI T seg_sum(T v, const SegRun<T>& r) {
    v = fma(dpp_down<1>(v), r.m1, v);
    v = fma(dpp_down<2>(v), r.m2, v);
    v = fma(dpp_down<4>(v), r.m4, v);
    v = fma(dpp_down<8>(v), r.m8, v);
    return v;
}

// seg_sum of NV values at once.  In FP32 each Hillis-Steele step is a single
// v_fmac_f32 with the DPP row shift on its first source (the compiler emits a
// separate v_mov_b32_dpp for every step instead): lane i adds m_d * c of lane
// i + d to its own c, and lanes whose source lies past their 16-lane row read
// 0 (bound_ctrl).  DPP reads of a VGPR need 2 wait states after the VALU
// write of that VGPR, and 5 after a VALU write of EXEC: the block opens with
// s_nop 4 (instructions before it may have written the values or EXEC), and
// the values are interleaved so that each is read again 3 (NV = 4) or 2
// (NV = 3) instructions after it was written.
template <typename T, int NV>
DMPM_DI void seg_sums(T* c, const SegRun<T>& r) {
#pragma unroll
    for (int q = 0; q < NV; ++q) c[q] = seg_sum(c[q], r);
}

#define DMPM_FMAC_DPP(dst, m, shift) \
    "v_fmac_f32_dpp " dst ", " dst ", " m " row_shl:" shift " row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"

template <>
DMPM_DI void seg_sums<float, 4>(float* c, const SegRun<float>& r) {
    asm volatile("s_nop 4\n\t"
                 DMPM_FMAC_DPP("%0", "%4", "1") DMPM_FMAC_DPP("%1", "%4", "1")
                 DMPM_FMAC_DPP("%2", "%4", "1") DMPM_FMAC_DPP("%3", "%4", "1")
                 DMPM_FMAC_DPP("%0", "%5", "2") DMPM_FMAC_DPP("%1", "%5", "2")
                 DMPM_FMAC_DPP("%2", "%5", "2") DMPM_FMAC_DPP("%3", "%5", "2")
                 DMPM_FMAC_DPP("%0", "%6", "4") DMPM_FMAC_DPP("%1", "%6", "4")
                 DMPM_FMAC_DPP("%2", "%6", "4") DMPM_FMAC_DPP("%3", "%6", "4")
                 DMPM_FMAC_DPP("%0", "%7", "8") DMPM_FMAC_DPP("%1", "%7", "8")
                 DMPM_FMAC_DPP("%2", "%7", "8") DMPM_FMAC_DPP("%3", "%7", "8")
                 : "+v"(c[0]), "+v"(c[1]), "+v"(c[2]), "+v"(c[3])
                 : "v"(r.m1), "v"(r.m2), "v"(r.m4), "v"(r.m8));
}

template <>
DMPM_DI void seg_sums<float, 3>(float* c, const SegRun<float>& r) {
    asm volatile("s_nop 4\n\t"
                 DMPM_FMAC_DPP("%0", "%3", "1") DMPM_FMAC_DPP("%1", "%3", "1") DMPM_FMAC_DPP("%2", "%3", "1")
                 DMPM_FMAC_DPP("%0", "%4", "2") DMPM_FMAC_DPP("%1", "%4", "2") DMPM_FMAC_DPP("%2", "%4", "2")
                 DMPM_FMAC_DPP("%0", "%5", "4") DMPM_FMAC_DPP("%1", "%5", "4") DMPM_FMAC_DPP("%2", "%5", "4")
                 DMPM_FMAC_DPP("%0", "%6", "8") DMPM_FMAC_DPP("%1", "%6", "8") DMPM_FMAC_DPP("%2", "%6", "8")
                 : "+v"(c[0]), "+v"(c[1]), "+v"(c[2])
                 : "v"(r.m1), "v"(r.m2), "v"(r.m4), "v"(r.m8));
}
#undef DMPM_FMAC_DPP

// Window-relative index of stencil base node b (block_window guarantees that
// every stencil node of a valid lane lies inside the window).
DMPM_DI int window_base(const Win& w, const int* b) {
    return (b[0] - w.lo0) * w.s1 + (b[1] - w.lo1) * w.s2 + (b[2] - w.lo2);
}

// Scatter NV values per stencil node: into the block's LDS window (NV Acc
// arrays of a.cap) or, when the window did not fit, straight into the grid
// dst (kNode values per node; 1 when NV = 1) with global atomics.  val(ia, ib, ic, c) writes the lane's
// contributions to node b + (ia, ib, ic); lanes without a particle must
// produce finite values (they never issue).  Run sums merge the lanes that
// share a base cell first.  Nodes outside the grid go to window slots that
// the flush drops, or are skipped on the global path.
template <typename T, int ORDER, int NV, typename Val>
DMPM_DI void scatter_nodes(const KP<T>& a, const Win& win, Acc* sm, const Sten<T, ORDER>& S, bool valid,
                           T* dst, Val&& val) {
    constexpr int KS = Spline<ORDER>::K;
    const SegRun<T> run = seg_run<T>(valid, S.b);
    const bool issue = valid && run.head;
    if (win.lds) {
        const int e2 = win.s2, e12 = win.s1;
        const int lb = valid ? window_base(win, S.b) : 0;
#pragma unroll
        for (int ia = 0; ia < KS; ++ia) {
#pragma unroll
            for (int ib = 0; ib < KS; ++ib) {
                const int row = lb + ia * e12 + ib * e2;
#pragma unroll
                for (int ic = 0; ic < KS; ++ic) {
                    T c[NV];
                    val(ia, ib, ic, c);
                    seg_sums<T, NV>(c, run);
                    if (issue) {
#pragma unroll
                        for (int q = 0; q < NV; ++q) atomicAdd(&sm[q * a.cap + row + ic], (Acc)c[q]);
                    }
                }
            }
        }
    } else {
#pragma unroll
        for (int ia = 0; ia < KS; ++ia) {
#pragma unroll
            for (int ib = 0; ib < KS; ++ib) {
#pragma unroll
                for (int ic = 0; ic < KS; ++ic) {
                    T c[NV];
                    val(ia, ib, ic, c);
                    seg_sums<T, NV>(c, run);
                    const int gi = S.b[0] + ia, gj = S.b[1] + ib, gk = S.b[2] + ic;
                    if (!issue || !in_grid(a, gi, gj, gk)) continue;
                    const int e = (int)gindex(a, gi, gj, gk) * (NV == 1 ? 1 : kNode);
#pragma unroll
                    for (int q = 0; q < NV; ++q) atomic_add_at(dst, e + q, c[q]);
                }
            }
        }
    }
}

// Scatter one particle's m and p = m v + A (x_i - x_p) into the window (or
// the grid).  Called by every lane; lanes without a particle pass m = 0,
// A = 0, mv = 0.
template <typename T, int ORDER, bool MASS_ONLY>
DMPM_DI void p2g_scatter(const KP<T>& a, const Win& win, Acc* sm, const Sten<T, ORDER>& S, bool valid, T m,
                         const m3<T>& A, const v3<T>& mv, T* dst) {
    scatter_nodes<T, ORDER, MASS_ONLY ? 1 : 4>(a, win, sm, S, valid, dst, [&](int ia, int ib, int ic, T* c) {
        const T d0 = S.u[0][ia] * a.dx, d1 = S.u[1][ib] * a.dx, d2 = S.u[2][ic] * a.dx;
        const T wt = S.w[0][ia] * S.w[1][ib] * S.w[2][ic];
        c[0] = wt * m;
        if (!MASS_ONLY) {
            c[1] = wt * (mv.a[0] + A.a[0] * d0 + A.a[1] * d1 + A.a[2] * d2);
            c[2] = wt * (mv.a[1] + A.a[3] * d0 + A.a[4] * d1 + A.a[5] * d2);
            c[3] = wt * (mv.a[2] + A.a[6] * d0 + A.a[7] * d1 + A.a[8] * d2);
        }
    });
}

template <typename T, int NA, typename U>
DMPM_DI void window_zero(const KP<T>& a, const Win& win, U* sm) {
    for (int t = threadIdx.x; t < win.vol; t += blockDim.x)
#pragma unroll
        for (int q = 0; q < NA; ++q) sm[q * a.cap + t] = U(0);
}

// Flush a scatter window (NV Acc arrays of a.cap) to the grid dst (kNode
// values per node; 1 when NV = 1): one atomic per nonzero node value.  Lane u
// takes value u % NV of window slot u / NV, so that the lanes of an atomic
// instruction cover a run of whole nodes, consecutive in memory along a
// window row.
template <typename T, int NV>
DMPM_DI void flush_window(const KP<T>& a, const Win& win, const Acc* sm, T* dst) {
    constexpr int NS = NV == 1 ? 1 : kNode;
    for (int u = threadIdx.x; u < NV * win.vol; u += blockDim.x) {
        const int t = u / NV, q = u - t * NV;
        int gi, gj, gk;
        if (!window_node(win, t, gi, gj, gk) || !in_grid(a, gi, gj, gk)) continue;
        const T v = (T)sm[q * a.cap + t];
        if (v != T(0)) atomic_add_at(dst, (int)gindex(a, gi, gj, gk) * NS + q, v);
    }
}

// P2G of layer `st` into the grid.
template <typename T, int ORDER, int MODEL, bool MASS_ONLY>
DMPM_DI void p2g_block(const KP<T>& a, int lblock, double* smem_d) {
    constexpr int KS = Spline<ORDER>::K;
    constexpr int NA = MASS_ONLY ? 1 : 4;
    Acc* sm = reinterpret_cast<Acc*>(smem_d);
    int i0, i1;
    block_range(a, lblock, i0, i1);
    const int i = i0 + (int)threadIdx.x;
    const bool valid = i < i1;
    Sten<T, ORDER> S;
    v3<T> x = vzero<T>();
    if (valid) x = load_v3(a.st, a.ld, PX, i);
    S.init(a, x);  // lanes without a particle get a finite stencil at x = 0
    const Win win = block_window<KS>(valid, S.b[0], S.b[1], S.b[2], a.cap);
    count_window(a, win);
    if (win.lds) window_zero<T, NA>(a, win, sm);  // Acc arrays
    T m = T(0);
    m3<T> A = mzero<T>();
    v3<T> mv = vzero<T>();
    if (valid) {
        m = load_s(a.stat, a.ld, SM, i);
        if (!MASS_ONLY) {
            m3<T> F = load_m3(a.st, a.ld, PF, i);
            if (a.ctl) F = madd(F, load_m3(a.ctl, a.ld, 0, i));
            p2g_terms<T, ORDER, MODEL>(a, i, load_v3(a.st, a.ld, PV, i), F, load_m3(a.st, a.ld, PC, i), m, A, mv);
        }
    }
    __syncthreads();
    p2g_scatter<T, ORDER, MASS_ONLY>(a, win, sm, S, valid, m, A, mv, a.gm);
    if (win.lds && win.any) {
        __syncthreads();
        flush_window<T, NA>(a, win, sm, a.gm);
    }
}

template <typename T, int ORDER, int MODEL, bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_p2g(KP<T> a) {
    extern __shared__ double smem_d[];
    p2g_block<T, ORDER, MODEL, MASS_ONLY>(a, logical_block(), smem_d);
}

template <typename T, int ORDER, int MODEL, bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_p2g_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[inst];
    p2g_block<T, ORDER, MODEL, MASS_ONLY>(a, lblock, smem_d);
}

// Window slots staged per thread and pass: a thread takes slots t,
// t + B, ..., t + (kStage - 1) B (B = blockDim.x) and issues every load of
// them before it uses any, so a pass costs one memory round trip.  Slots
// past the window or off the grid load node 0 and are zeroed.
constexpr int kStage = 4;
constexpr int kStageAdj = 2;  // the adjoint loads eight values per slot: m, p and dL/dv of the node

// Stage the window's grid velocities in LDS (3 arrays of cap).
template <typename T>
DMPM_DI void stage_velocity(const KP<T>& a, const Win& win, T* sv) {
    const int B = blockDim.x;
    for (int t0 = threadIdx.x; t0 < win.vol; t0 += kStage * B) {
        T n[kStage][kNode];
        bool live[kStage];
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            int gi, gj, gk;
            const int t = t0 + u * B;
            const bool node = t < win.vol && window_node(win, t, gi, gj, gk) && in_grid(a, gi, gj, gk);
            load_node(a.gm, node ? (int)gindex(a, gi, gj, gk) : 0, n[u]);
            live[u] = node && !is_wall(a, gi, gj, gk);
        }
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
            const int t = t0 + u * B;
            T v0, v1, v2;
            velocity_of(a, live[u], n[u][0], n[u][1], n[u][2], n[u][3], v0, v1, v2);
            if (t < win.vol) {
                sv[t] = v0;
                sv[a.cap + t] = v1;
                sv[2 * a.cap + t] = v2;
            }
        }
    }
}

// G2P gather (ForwardSimulation.cpp:89-123): v' = sum w v_i,
// C' = D^{-1} sum w v_i (x_i - x_p)^T.
template <typename T, int ORDER>
DMPM_DI void g2p_gather(const KP<T>& a, const Win& win, const T* sv, const Sten<T, ORDER>& S, v3<T>& vn,
                        m3<T>& Craw) {
    constexpr int KS = Spline<ORDER>::K;
    vn = vzero<T>();
    Craw = mzero<T>();
    auto add = [&](int ia, int ib, int ic, T v0, T v1, T v2) {
        const T d0 = S.u[0][ia] * a.dx, d1 = S.u[1][ib] * a.dx, d2 = S.u[2][ic] * a.dx;
        const T wt = S.w[0][ia] * S.w[1][ib] * S.w[2][ic];
        const T w0 = wt * v0, w1 = wt * v1, w2 = wt * v2;
        vn.a[0] += w0; vn.a[1] += w1; vn.a[2] += w2;
        Craw.a[0] += w0 * d0; Craw.a[1] += w0 * d1; Craw.a[2] += w0 * d2;
        Craw.a[3] += w1 * d0; Craw.a[4] += w1 * d1; Craw.a[5] += w1 * d2;
        Craw.a[6] += w2 * d0; Craw.a[7] += w2 * d1; Craw.a[8] += w2 * d2;
    };
    if (win.lds) {
        // Window nodes outside the grid were staged with zero velocity.  The
        // weight w0 w1 w2 and the offset (d0, d1, d2) are separable per axis,
        // so the sums run over z, then y, then x: per (x, y) column
        // s = sum_z w2 v and t = sum_z w2 d2 v, per x plane
        // S = sum_y w1 s, U = sum_y w1 d1 s, T = sum_y w1 t, and finally
        // v' = sum_x w0 S, C' columns sum_x w0 d0 S, sum_x w0 U, sum_x w0 T
        // (about 2/3 of the per-node arithmetic).
        const int e2 = win.s2, e12 = win.s1;
        const int lb = window_base(win, S.b);
        T w2d[KS];
#pragma unroll
        for (int ic = 0; ic < KS; ++ic) w2d[ic] = S.w[2][ic] * (S.u[2][ic] * a.dx);
#pragma unroll
        for (int ia = 0; ia < KS; ++ia) {
            T Sa[3] = {T(0), T(0), T(0)}, Ua[3] = {T(0), T(0), T(0)}, Ta[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int ib = 0; ib < KS; ++ib) {
                const int row = lb + ia * e12 + ib * e2;
                T sc[3] = {T(0), T(0), T(0)}, tc[3] = {T(0), T(0), T(0)};
#pragma unroll
                for (int ic = 0; ic < KS; ++ic) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const T v = sv[q * a.cap + row + ic];
                        sc[q] += S.w[2][ic] * v;
                        tc[q] += w2d[ic] * v;
                    }
                }
                const T w1 = S.w[1][ib], w1d = w1 * (S.u[1][ib] * a.dx);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    Sa[q] += w1 * sc[q];
                    Ua[q] += w1d * sc[q];
                    Ta[q] += w1 * tc[q];
                }
            }
            const T w0 = S.w[0][ia], w0d = w0 * (S.u[0][ia] * a.dx);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                vn.a[q] += w0 * Sa[q];
                Craw.a[3 * q + 0] += w0d * Sa[q];
                Craw.a[3 * q + 1] += w0 * Ua[q];
                Craw.a[3 * q + 2] += w0 * Ta[q];
            }
        }
    } else {
#pragma unroll
        for (int ia = 0; ia < KS; ++ia)
#pragma unroll
            for (int ib = 0; ib < KS; ++ib)
#pragma unroll
                for (int ic = 0; ic < KS; ++ic) {
                    const int gi = S.b[0] + ia, gj = S.b[1] + ib, gk = S.b[2] + ic;
                    if (!in_grid(a, gi, gj, gk)) continue;
                    T v0, v1, v2;
                    node_velocity(a, gindex(a, gi, gj, gk), gi, gj, gk, v0, v1, v2);
                    add(ia, ib, ic, v0, v1, v2);
                }
    }
    Craw = mscale(Spline<ORDER>::dinv(a.dx), Craw);
}

// P_op_2 (ForwardSimulation.cpp:131-157): trust region on C', F update with
// smoothing, x update.  Writes layer t+1 and returns its x, v, F, C.
template <typename T>
DMPM_DI void particle_update(const KP<T>& a, int i, const v3<T>& x, const m3<T>& F, const m3<T>& Ft,
                             const v3<T>& vn, const m3<T>& Craw, v3<T>& xn, m3<T>& Fn, m3<T>& Cn) {
    const Trust<T> tr = trust_region(a, Ft, Craw);
    Cn = tr.state ? mscale(tr.s, Craw) : Craw;
    m3<T> A1 = mscale(a.dt, Cn);
    A1.a[0] += T(1); A1.a[4] += T(1); A1.a[8] += T(1);
    const m3<T> Fu = mmul(A1, Ft);
#pragma unroll
    for (int q = 0; q < 9; ++q) Fn.a[q] = (T(1) - a.sm) * Fu.a[q] + a.sm * F.a[q];
    xn = {{x.a[0] + a.dt * vn.a[0], x.a[1] + a.dt * vn.a[1], x.a[2] + a.dt * vn.a[2]}};
    const int io = a.opos ? a.opos[i] : i;
    store_v3(a.out, a.ld, PX, io, xn);
    store_v3(a.out, a.ld, PV, io, vn);
    store_m3(a.out, a.ld, PF, io, Fn);
    store_m3(a.out, a.ld, PC, io, Cn);
}

template <typename T>
DMPM_DI void particle_update(const KP<T>& a, int i, const v3<T>& x, const v3<T>& vn, const m3<T>& Craw,
                             v3<T>& xn, m3<T>& Fn, m3<T>& Cn) {
    const m3<T> F = load_m3(a.st, a.ld, PF, i);
    const m3<T> Ft = a.ctl ? madd(F, load_m3(a.ctl, a.ld, 0, i)) : F;
    particle_update(a, i, x, F, Ft, vn, Craw, xn, Fn, Cn);
}

// G2P + P_op_2 (ForwardSimulation.cpp:89-157): writes layer t+1.
template <typename T, int ORDER>
DMPM_DI void g2p_block(const KP<T>& a, int lblock, double* smem_d) {
    constexpr int KS = Spline<ORDER>::K;
    T* sv = reinterpret_cast<T*>(smem_d);
    int i0, i1;
    block_range(a, lblock, i0, i1);
    const int i = i0 + (int)threadIdx.x;
    const bool valid = i < i1;
    Sten<T, ORDER> S;
    v3<T> x = vzero<T>();
    if (valid) x = load_v3(a.st, a.ld, PX, i);
    S.init(a, x);  // lanes without a particle get a finite stencil at x = 0
    const Win win = block_window<KS>(valid, S.b[0], S.b[1], S.b[2], a.cap);
    if (win.lds) {
        stage_velocity(a, win, sv);
        __syncthreads();
    }
    if (!valid) return;
    v3<T> vn, xn;
    m3<T> Craw, Fn, Cn;
    g2p_gather<T, ORDER>(a, win, sv, S, vn, Craw);
    particle_update(a, i, x, vn, Craw, xn, Fn, Cn);
}

template <typename T, int ORDER>
__global__ void __launch_bounds__(256) k_g2p(KP<T> a) {
    extern __shared__ double smem_d[];
    g2p_block<T, ORDER>(a, logical_block(), smem_d);
}

template <typename T, int ORDER>
__global__ void __launch_bounds__(256) k_g2p_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[inst];
    g2p_block<T, ORDER>(a, lblock, smem_d);
}

// One fused forward step: G2P of step t from grid t (a.gm), the
// particle update writing layer t+1, then the stress of layer t+1 (with its
// control a.ctl1) and its P2G into grid t+1 (a.gm1, zeroed by the
// caller).  Layer t+1 is never re-read from HBM for its own P2G.
template <typename T, int ORDER, int MODEL>
DMPM_DI void step_block(const KP<T>& a, int lblock, double* smem_d) {
    constexpr int KS = Spline<ORDER>::K;
    T* sm = reinterpret_cast<T*>(smem_d);     // staged velocities
    Acc* sacc = reinterpret_cast<Acc*>(smem_d);  // then the P2G accumulators
    int i0, i1;
    block_range(a, lblock, i0, i1);
    const int i = i0 + (int)threadIdx.x;
    const bool valid = i < i1;
    // Every per-particle input is loaded up front, so that the memory latency
    // overlaps the window reduction and the staging of grid velocities
    // (loads are not moved across the barriers by the compiler).
    Sten<T, ORDER> S;
    v3<T> x = vzero<T>();
    m3<T> F = mzero<T>(), Ft = mzero<T>(), ctl1 = mzero<T>();
    T m = T(0), vol = T(0), lam = T(0), mu = T(0);
    if (valid) {
        x = load_v3(a.st, a.ld, PX, i);
        F = load_m3(a.st, a.ld, PF, i);
        Ft = a.ctl ? madd(F, load_m3(a.ctl, a.ld, 0, i)) : F;
        if (a.ctl1) ctl1 = load_m3(a.ctl1, a.ld, 0, i);
        m = load_s(a.stat, a.ld, SM, i);
        vol = load_s(a.stat, a.ld, SVOL, i);
        lam = load_s(a.stat, a.ld, SLAM, i);
        mu = load_s(a.stat, a.ld, SMU, i);
    }
    S.init(a, x);  // lanes without a particle get a finite stencil at x = 0
    const Win win = block_window<KS>(valid, S.b[0], S.b[1], S.b[2], a.cap);
    if (win.lds) {
        stage_velocity(a, win, sm);
        __syncthreads();
    }
    v3<T> vn = vzero<T>(), xn = vzero<T>();
    m3<T> Craw, Fn = mzero<T>(), Cn = mzero<T>();
    if (valid) {
        g2p_gather<T, ORDER>(a, win, sm, S, vn, Craw);
        particle_update(a, i, x, F, Ft, vn, Craw, xn, Fn, Cn);
    }
    // The next window's reduction overwrites the first one's table, and its
    // barrier (inside block_window) is the one after which the accumulators,
    // which alias the staged velocities, are cleared.  With the staging barrier
    // above, every wave has read the first table already; without it (the
    // window did not fit), a barrier here keeps the tables apart.
    if (!win.lds) __syncthreads();
    Sten<T, ORDER> S1;
    S1.init(a, xn);
    const Win win1 = block_window<KS>(valid, S1.b[0], S1.b[1], S1.b[2], a.cap);
    count_window(a, win1);
    if (win1.lds) window_zero<T, 4>(a, win1, sacc);
    m3<T> A = mzero<T>();
    v3<T> mv = vzero<T>();
    if (valid) {
        const m3<T> Ft1 = a.ctl1 ? madd(Fn, ctl1) : Fn;
        p2g_terms<T, ORDER, MODEL>(a, vol, lam, mu, vn, Ft1, Cn, m, A, mv);
    }
    __syncthreads();
    p2g_scatter<T, ORDER, false>(a, win1, sacc, S1, valid, m, A, mv, a.gm1);
    if (win1.lds && win1.any) {
        __syncthreads();
        flush_window<T, 4>(a, win1, sacc, a.gm1);
    }
}

template <typename T, int ORDER, int MODEL>
__global__ void __launch_bounds__(256) k_step(KP<T> a) {
    extern __shared__ double smem_d[];
    const int lb = logical_block();
    step_block<T, ORDER, MODEL>(a, lb, smem_d);
    clear_share(a, lb);
}

template <typename T, int ORDER, int MODEL>
__global__ void __launch_bounds__(256) k_step_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[inst];
    step_block<T, ORDER, MODEL>(a, lblock, smem_d);
    clear_share(a, lblock);
}

// The fixed-corotated FP32 adjoint parks the polar factors R (9 values) and
// the upper triangle of S (6) of F + dFc in LDS through its node loop, after
// the 12 phase-1 terms: its blocks hold three waves per SIMD, so the LDS has
// room for them.  (FP64 would need twice the room; it recomputes instead.)
template <typename T, int MODEL>
__host__ __device__ constexpr bool park_polar() { return MODEL == 0 && sizeof(T) == 4; }
template <typename T>
__host__ __device__ constexpr int park_values(int model) { return model == 0 && sizeof(T) == 4 ? 27 : 12; }
__device__ constexpr int kSymUpper[6] = {0, 1, 2, 4, 5, 8};

// Phase 1 of the adjoint (BackPropagation.cpp:22-32).  REF reproduces the
// reference's formulas; otherwise the exact VJP of P_op_2 including the
// trust region and the smoothing blend.
template <typename T, bool REF>
struct Phase1 {
    v3<T> gvtot;  // dL/dv' (including dt dL/dx')
    m3<T> gCraw;  // dL/dC' before the trust region
    m3<T> gFt;    // dL/d(F+dFc) from the F update
    m3<T> gFdir;  // dL/dF through the smoothing blend (F only, not dFc)
};

// Adjoint carry of layer t+1 for one particle: dL/dx', dL/dv', dL/dF', dL/dC'.
template <typename T>
struct Carry {
    v3<T> gx, gv;
    m3<T> gF, gC;
};

template <typename T>
DMPM_DI Carry<T> load_carry(const KP<T>& a, int i) {
    Carry<T> c;
    c.gx = load_v3(a.adj1, a.ld, PX, i);
    c.gv = load_v3(a.adj1, a.ld, PV, i);
    c.gF = load_m3(a.adj1, a.ld, PF, i);
    c.gC = load_m3(a.adj1, a.ld, PC, i);
    return c;
}

template <typename T, bool REF>
DMPM_DI Phase1<T, REF> phase1(const KP<T>& a, const Carry<T>& cy, const m3<T>& Ft, const m3<T>& Craw,
                              const Trust<T>& tr) {
    Phase1<T, REF> r;
    const v3<T>& gx1 = cy.gx;
    const v3<T>& gv1 = cy.gv;
    const m3<T>& gF1 = cy.gF;
    const m3<T>& gC1 = cy.gC;
    const m3<T> Cn = tr.state ? mscale(tr.s, Craw) : Craw;
    m3<T> A1 = mscale(a.dt, Cn);
    A1.a[0] += T(1); A1.a[4] += T(1); A1.a[8] += T(1);
#pragma unroll
    for (int q = 0; q < 3; ++q) r.gvtot.a[q] = a.dt * gx1.a[q] + gv1.a[q];
    const m3<T> AtG = mtmul(A1, gF1);
    const m3<T> GFt = mmult(gF1, Ft);  // gF1 Ft^T
    if (REF) {
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            r.gFt.a[q] = (T(1) - a.sm) * AtG.a[q] + a.sm * gF1.a[q];
            r.gFdir.a[q] = T(0);
            r.gCraw.a[q] = a.dt * GFt.a[q] + gC1.a[q];
        }
    } else {
        m3<T> gCfin;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            r.gFt.a[q] = (T(1) - a.sm) * AtG.a[q];
            r.gFdir.a[q] = a.sm * gF1.a[q];
            gCfin.a[q] = gC1.a[q] + (T(1) - a.sm) * a.dt * GFt.a[q];
        }
        if (tr.state == 0) {
            r.gCraw = gCfin;
        } else if (tr.state == 2) {
            r.gCraw = mzero<T>();
        } else {
            const T cc = mddot(gCfin, Craw);
            r.gCraw = mscale(tr.s, gCfin);
            const T trC = mtrace(Craw);
            if (a.dt * trC < T(-1e-6)) {
                const T d = -tr.s / trC * cc;
                r.gCraw.a[0] += d; r.gCraw.a[4] += d; r.gCraw.a[8] += d;
            }
            const m3<double> Ftd = mcast<double>(Ft);
            const double detF = mdet(Ftd);
            if (detF > 1e-6) {
                const double dsdJ = -(double)a.jmin / ((double)tr.J0 * (double)tr.J0) / (double)tr.denom;
                const m3<double> cf = mcof(Ftd);
                const double k = (double)cc * dsdJ;
#pragma unroll
                for (int q = 0; q < 9; ++q) r.gFt.a[q] += (T)(k * cf.a[q]);
            }
        }
    }
    return r;
}

template <typename T, bool REF>
DMPM_DI Phase1<T, REF> phase1(const KP<T>& a, int i, const m3<T>& Ft, const m3<T>& Craw, const Trust<T>& tr) {
    return phase1<T, REF>(a, load_carry(a, i), Ft, Craw, tr);
}

template <typename T, int ORDER, bool REF>
DMPM_DI void adj_scatter_core(const KP<T>& a, bool valid, const v3<T>& x, const m3<T>& Ft, m3<T> Craw,
                              const Carry<T>& cy, bool nxt, T* sv, Acc* sg);

// Adjoint phases 1 + 2: dL/dv_i += w gvtot + D^{-1} w gCraw (x_i - x_p)
// (BackPropagation.cpp:36-55), accumulated in an LDS window.
template <typename T, int ORDER, bool REF>
DMPM_DI void adj_scatter_block(const KP<T>& a, int lblock, double* smem_d) {
    // 3 velocity arrays, then 3 accumulators; reading layer t+1 (no G2P
    // recompute, so no velocities staged), the accumulators alone
    const bool nxt = reads_next<T, REF>(a);
    T* sv = reinterpret_cast<T*>(smem_d);
    Acc* sg = reinterpret_cast<Acc*>(smem_d + (nxt ? 0 : (3 * a.cap * sizeof(T) + 7) / 8));
    int i0, i1;
    block_range(a, lblock, i0, i1);
    const int i = i0 + (int)threadIdx.x;
    const bool valid = i < i1;
    // Every per-particle input is loaded up front, so that the memory latency
    // overlaps the window reduction and the clearing of the accumulators
    // (loads are not moved across the barriers by the compiler).
    v3<T> x = vzero<T>();
    m3<T> Ft = mzero<T>(), Craw = mzero<T>();
    Carry<T> cy;
    if (valid) {
        x = load_v3(a.st, a.ld, PX, i);
        Ft = load_m3(a.st, a.ld, PF, i);
        if (a.ctl) Ft = madd(Ft, load_m3(a.ctl, a.ld, 0, i));
        if (nxt) Craw = load_m3(a.nxt, a.ld, PC, i);
        cy = load_carry(a, i);
    }
    adj_scatter_core<T, ORDER, REF>(a, valid, x, Ft, Craw, cy, nxt, sv, sg);
}

// Phases 1 + 2 of one step for the block's particles, given their inputs:
// x, F + dFc of layer t, C' (of layer t+1 when nxt, else recomputed by a G2P
// from the grid staged in sv) and the carry of layer t+1.  Every thread of
// the block must call this.
template <typename T, int ORDER, bool REF>
DMPM_DI void adj_scatter_core(const KP<T>& a, bool valid, const v3<T>& x, const m3<T>& Ft, m3<T> Craw,
                              const Carry<T>& cy, bool nxt, T* sv, Acc* sg) {
    constexpr int KS = Spline<ORDER>::K;
    Sten<T, ORDER> S;
    S.init(a, x);  // lanes without a particle get a finite stencil at x = 0
    const Win win = block_window<KS>(valid, S.b[0], S.b[1], S.b[2], a.cap);
    count_window(a, win);
    if (win.lds) {
        if (!nxt) stage_velocity(a, win, sv);
        for (int t = threadIdx.x; t < win.vol; t += blockDim.x) {
            sg[t] = Acc(0);
            sg[a.cap + t] = Acc(0);
            sg[2 * a.cap + t] = Acc(0);
        }
        __syncthreads();
    }
    v3<T> gvtot = vzero<T>();
    m3<T> Gc = mzero<T>();
    if (valid) {
        Trust<T> tr;
        if (nxt) {
            tr = trust_none<T>();
        } else {
            v3<T> vn;
            g2p_gather<T, ORDER>(a, win, sv, S, vn, Craw);
            tr = trust_region(a, Ft, Craw);
        }
        const Phase1<T, REF> p1 = phase1<T, REF>(a, cy, Ft, Craw, tr);
        gvtot = p1.gvtot;
        Gc = mscale(Spline<ORDER>::dinv(a.dx), p1.gCraw);
    }
    // dL/dv_i += w (dL/dv' + D^{-1} dL/dC' (x_i - x_p)); lanes without a
    // particle scatter zeros
    T* const dst = a.ggv;
    scatter_nodes<T, ORDER, 3>(a, win, sg, S, valid, dst, [&](int ia, int ib, int ic, T* c) {
        const T d0 = S.u[0][ia] * a.dx, d1 = S.u[1][ib] * a.dx, d2 = S.u[2][ic] * a.dx;
        const T wt = S.w[0][ia] * S.w[1][ib] * S.w[2][ic];
        c[0] = wt * (gvtot.a[0] + Gc.a[0] * d0 + Gc.a[1] * d1 + Gc.a[2] * d2);
        c[1] = wt * (gvtot.a[1] + Gc.a[3] * d0 + Gc.a[4] * d1 + Gc.a[5] * d2);
        c[2] = wt * (gvtot.a[2] + Gc.a[6] * d0 + Gc.a[7] * d1 + Gc.a[8] * d2);
    });
    if (win.lds && win.any) {
        __syncthreads();
        flush_window<T, 3>(a, win, sg, a.ggv);
    }
}

template <typename T, int ORDER, bool REF>
__global__ void __launch_bounds__(256) k_adj_scatter(KP<T> a) {
    extern __shared__ double smem_d[];
    adj_scatter_block<T, ORDER, REF>(a, logical_block(), smem_d);
}

template <typename T, int ORDER, bool REF>
__global__ void __launch_bounds__(256) k_adj_scatter_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[inst];
    adj_scatter_block<T, ORDER, REF>(a, lblock, smem_d);
}

// Grid adjoint (BackPropagation.cpp:59-72) of a node whose mass, momentum
// q and dL/dv_i g are loaded: h = dL/dp_i, gm = dL/dm_i.  Exact: dm uses
// v_i - dt f_ext; REF uses v_i (as written).  `live` is false for walls and
// nodes off the grid; selects rather than branches keep the loads ahead.
template <typename T, bool REF>
DMPM_DI void adjoint_of(const KP<T>& a, bool live, T m, const T* q, const T* gv, T* v, T* h, T& gmass) {
    const T inv = T(1) / m;
    live = live && m > T(1e-12);
    const T p0 = q[0] * inv, p1 = q[1] * inv, p2 = q[2] * inv;
    const T u0 = p0 + a.dt * a.f0, u1 = p1 + a.dt * a.f1, u2 = p2 + a.dt * a.f2;
    const T vd = REF ? (u0 * gv[0] + u1 * gv[1] + u2 * gv[2]) : (p0 * gv[0] + p1 * gv[1] + p2 * gv[2]);
    v[0] = live ? u0 : T(0);
    v[1] = live ? u1 : T(0);
    v[2] = live ? u2 : T(0);
    h[0] = live ? gv[0] * inv : T(0);
    h[1] = live ? gv[1] * inv : T(0);
    h[2] = live ? gv[2] * inv : T(0);
    gmass = live ? -inv * vd : T(0);
}

// Grid adjoint (BackPropagation.cpp:59-72) for one node: h = dL/dp_i,
// gm = dL/dm_i.  Exact: dm uses v_i - dt f_ext; REF uses v_i (as written).
template <typename T, bool REF>
DMPM_DI void node_adjoint(const KP<T>& a, long long g, int i, int j, int k, T* v, T* h, T& gmass) {
    T n[kNode], gv[kNode];
    load_node(a.gm, (int)g, n);
    load_node(a.ggv, (int)g, gv);
    adjoint_of<T, REF>(a, !is_wall(a, i, j, k), n[0], n + 1, gv, v, h, gmass);
}

// Adjoint phases 3 + 4 and the stress VJP (BackPropagation.cpp:59-143).
// The window holds per node v_i (3), h_i = dL/dp_i (3), dL/dm_i (1).
// FUSED: then phases 1 + 2 of the previous step (t-1, parameters *b) for
// the same particles, with the carry of layer t just computed kept in
// registers and C' of step t-1 = C of layer t (the caller fuses only when
// layer t is the forward step of layer t-1 and the adjoint reads C' from
// it); b->ggv must be a different grid from a.ggv.
template <typename T, int ORDER, int MODEL, bool REF, bool FUSED>
DMPM_DI void adj_gather_block(const KP<T>& a, const KP<T>* b, int lblock, double* smem_d) {
    constexpr int KS = Spline<ORDER>::K;
    T* sv = reinterpret_cast<T*>(smem_d);  // v (3 arrays), h (3), gm (1)
    int i0, i1;
    block_range(a, lblock, i0, i1);
    const int i = i0 + (int)threadIdx.x;
    const bool valid = i < i1;
    const bool nxt = reads_next<T, REF>(a);
    // Per-particle inputs are loaded up front, so that the memory latency
    // overlaps the window reduction and the staging of the node values
    // (loads are not moved across the barriers by the compiler).
    Sten<T, ORDER> S;
    v3<T> x = vzero<T>(), vn = vzero<T>();
    m3<T> F = mzero<T>(), Ft = mzero<T>(), Craw = mzero<T>();
    Carry<T> cy;
    if (valid) {
        x = load_v3(a.st, a.ld, PX, i);
        F = load_m3(a.st, a.ld, PF, i);
        Ft = a.ctl ? madd(F, load_m3(a.ctl, a.ld, 0, i)) : F;
        if (nxt) {
            vn = load_v3(a.nxt, a.ld, PV, i);
            Craw = load_m3(a.nxt, a.ld, PC, i);
        }
        cy = load_carry(a, i);
    }
    S.init(a, x);  // lanes without a particle get a finite stencil at x = 0
    const Win win = block_window<KS>(valid, S.b[0], S.b[1], S.b[2], a.cap);
    if (win.lds) {
        // kStageAdj slots per thread and pass, the eight values of each
        // loaded before any is used (slots past the window or off the grid
        // load node 0 and are zeroed)
        constexpr int U = kStageAdj;
        const int B = blockDim.x;
        for (int t0 = threadIdx.x; t0 < win.vol; t0 += U * B) {
            T n[U][kNode], gv[U][kNode];
            bool live[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int gi, gj, gk;
                const int t = t0 + u * B;
                const bool node = t < win.vol && window_node(win, t, gi, gj, gk) && in_grid(a, gi, gj, gk);
                const int g = node ? (int)gindex(a, gi, gj, gk) : 0;
                load_node(a.gm, g, n[u]);
                load_node(a.ggv, g, gv[u]);
                live[u] = node && !is_wall(a, gi, gj, gk);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int t = t0 + u * B;
                T v[3], h[3], gmass;
                adjoint_of<T, REF>(a, live[u], n[u][0], n[u] + 1, gv[u], v, h, gmass);
                if (t < win.vol) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        sv[c * a.cap + t] = v[c];
                        sv[(3 + c) * a.cap + t] = h[c];
                    }
                    sv[6 * a.cap + t] = gmass;
                }
            }
        }
        __syncthreads();
    }
    Carry<T> cyt;  // carry of layer t (the outputs)
    cyt.gx = cyt.gv = vzero<T>();
    cyt.gF = cyt.gC = mzero<T>();
    if (valid) {
        Trust<T> tr;
        if (nxt) {
            tr = trust_none<T>();
        } else {
            g2p_gather<T, ORDER>(a, win, sv, S, vn, Craw);
            tr = trust_region(a, Ft, Craw);
        }
        const Phase1<T, REF> p1 = phase1<T, REF>(a, cy, Ft, Craw, tr);

        const T m = load_s(a.stat, a.ld, SM, i), vol = load_s(a.stat, a.ld, SVOL, i);
        const T lam = load_s(a.stat, a.ld, SLAM, i), mu = load_s(a.stat, a.ld, SMU, i);
        const v3<T> vp = load_v3(a.st, a.ld, PV, i);
        const m3<T> C = load_m3(a.st, a.ld, PC, i);
        const T dinv = Spline<ORDER>::dinv(a.dx);
        const T k0 = dinv * a.dt * vol;
        // Gt = A_p = -k0 (P Ft^T + zeta tr(C) I) + m C
        m3<T> Gt;
        {
            m3<T> PFt;
            if constexpr (park_polar<T, MODEL>()) {
                // the polar factors of F + dFc are parked with the phase-1
                // terms below and used again by the stress gradient after
                // the node loop (one polar decomposition per particle, not two)
                const typename Model<0>::template Parts<T> fp = Model<0>::parts(Ft);
                PFt = mmult(Model<0>::pk1(fp, Ft, lam, mu), Ft);
                T* park = reinterpret_cast<T*>(reinterpret_cast<char*>(smem_d) + a.park);
                const int pt = threadIdx.x, pb = blockDim.x;
#pragma unroll
                for (int q = 0; q < 9; ++q) park[(12 + q) * pb + pt] = fp.R.a[q];
#pragma unroll
                for (int q = 0; q < 6; ++q) park[(21 + q) * pb + pt] = fp.S.a[kSymUpper[q]];
            } else {
                PFt = Model<MODEL>::kirchhoff(Ft, lam, mu);
            }
    #pragma unroll
            for (int q = 0; q < 9; ++q) Gt.a[q] = -k0 * PFt.a[q] + m * C.a[q];
            const T bulk = -k0 * a.zeta * mtrace(C);
            Gt.a[0] += bulk; Gt.a[4] += bulk; Gt.a[8] += bulk;
        }
        const T mom = REF ? m : m * (T(1) - a.dt * a.drag);
        const m3<T> Gc = mscale(dinv, p1.gCraw);
        const v3<T> mvp = {{mom * vp.a[0], mom * vp.a[1], mom * vp.a[2]}};
        const v3<T> gvtot = p1.gvtot;
        // What the code after the node loop needs from phase 1 is formed now
        // and parked in LDS for the loop's duration (12 values per thread,
        // lane-consecutive): the dL/dF terms other than the stress's,
        //   Q = (1-s) A'^T dL/dF' + s dL/dF'            (exact)
        //   Q = (1-s) A'^T dL/dF' + s dL/dF' + A'^T dL/dF'  (reference: added
        //       twice, BackPropagation.cpp:27-28 and :140),
        // and the dL/dx terms other than the gather's, dL/dx' - Gc^T v'.  The
        // carry, v' and C' are then not reloaded after the loop.
        {
            m3<T> Q = p1.gFt;
            if (REF) {
                const m3<T> Cn = tr.state ? mscale(tr.s, Craw) : Craw;
                m3<T> A1 = mscale(a.dt, Cn);
                A1.a[0] += T(1); A1.a[4] += T(1); A1.a[8] += T(1);
                Q = madd(Q, mtmul(A1, cy.gF));
            } else {
                Q = madd(Q, p1.gFdir);
            }
            const v3<T> Cv = mtv(Gc, vn);
            T* park = reinterpret_cast<T*>(reinterpret_cast<char*>(smem_d) + a.park);
            const int pt = threadIdx.x, pb = blockDim.x;
#pragma unroll
            for (int q = 0; q < 9; ++q) park[q * pb + pt] = Q.a[q];
#pragma unroll
            for (int q = 0; q < 3; ++q) park[(9 + q) * pb + pt] = cy.gx.a[q] - Cv.a[q];
            // dL/d dFc of the layer: the same terms less the smoothing
            // blend's in the exact adjoint; the stress term is added below
            if (a.gctl) store_m3(a.gctl, a.ld, 0, i, REF ? Q : p1.gFt);
        }

        // accumulators
        m3<T> gA = mzero<T>();   // sum w h dgp^T
        v3<T> sh = vzero<T>();   // sum w h
        v3<T> dxa = vzero<T>();  // sum s_i grad w
        const T invdx = T(-1) / a.dx;
        auto add = [&](int ia, int ib, int ic, const T* v, const T* h, T gmass) {
            const T d0 = S.u[0][ia] * a.dx, d1 = S.u[1][ib] * a.dx, d2 = S.u[2][ic] * a.dx;
            const T wt = S.w[0][ia] * S.w[1][ib] * S.w[2][ic];
            // grad_{x_p} w = -(1/dx) (N' N N, N N' N, N N N')
            const T gw0 = invdx * S.dw[0][ia] * S.w[1][ib] * S.w[2][ic];
            const T gw1 = invdx * S.w[0][ia] * S.dw[1][ib] * S.w[2][ic];
            const T gw2 = invdx * S.w[0][ia] * S.w[1][ib] * S.dw[2][ic];
            // q = mom v_p + Gt dgp ; (Gc dgp)
            const T Gd0 = Gt.a[0] * d0 + Gt.a[1] * d1 + Gt.a[2] * d2;
            const T Gd1 = Gt.a[3] * d0 + Gt.a[4] * d1 + Gt.a[5] * d2;
            const T Gd2 = Gt.a[6] * d0 + Gt.a[7] * d1 + Gt.a[8] * d2;
            const T Cd0 = Gc.a[0] * d0 + Gc.a[1] * d1 + Gc.a[2] * d2;
            const T Cd1 = Gc.a[3] * d0 + Gc.a[4] * d1 + Gc.a[5] * d2;
            const T Cd2 = Gc.a[6] * d0 + Gc.a[7] * d1 + Gc.a[8] * d2;
            const T s = m * gmass + (mvp.a[0] + Gd0) * h[0] + (mvp.a[1] + Gd1) * h[1] + (mvp.a[2] + Gd2) * h[2] +
                        v[0] * (gvtot.a[0] + Cd0) + v[1] * (gvtot.a[1] + Cd1) + v[2] * (gvtot.a[2] + Cd2);
            dxa.a[0] += s * gw0;
            dxa.a[1] += s * gw1;
            dxa.a[2] += s * gw2;
            const T wh0 = wt * h[0], wh1 = wt * h[1], wh2 = wt * h[2];
            sh.a[0] += wh0; sh.a[1] += wh1; sh.a[2] += wh2;
            gA.a[0] += wh0 * d0; gA.a[1] += wh0 * d1; gA.a[2] += wh0 * d2;
            gA.a[3] += wh1 * d0; gA.a[4] += wh1 * d1; gA.a[5] += wh1 * d2;
            gA.a[6] += wh2 * d0; gA.a[7] += wh2 * d1; gA.a[8] += wh2 * d2;
        };
        if (win.lds) {
            // Window nodes outside the grid were staged as zeros.  The same
            // sums as add(), grouped by axis: the node weight w0 w1 w2, its
            // gradient (-1/dx)(w0' w1 w2, w0 w1' w2, w0 w1 w2') and the
            // offsets d are separable, so each (x, y) column first sums over
            // z with the z factors -- s w2, s w2', h w2, h w2 d2 -- and the
            // x, y factors are applied once per column.  The column part of
            // Gt d and Gc d is also formed once per column.
            const int e2 = win.s2, e12 = win.s1;
            const int lb = window_base(win, S.b);
            T d2[KS], w2d[KS];
    #pragma unroll
            for (int ic = 0; ic < KS; ++ic) {
                d2[ic] = S.u[2][ic] * a.dx;
                w2d[ic] = S.w[2][ic] * d2[ic];
            }
    #pragma unroll
            for (int ia = 0; ia < KS; ++ia) {
                const T d0 = S.u[0][ia] * a.dx;
    #pragma unroll
                for (int ib = 0; ib < KS; ++ib) {
                    const T d1 = S.u[1][ib] * a.dx;
                    const int row = lb + ia * e12 + ib * e2;
                    T g01[3], c01[3];
    #pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        g01[q] = mvp.a[q] + Gt.a[3 * q] * d0 + Gt.a[3 * q + 1] * d1;
                        c01[q] = gvtot.a[q] + Gc.a[3 * q] * d0 + Gc.a[3 * q + 1] * d1;
                    }
                    T zs = T(0), zsd = T(0), zh[3] = {T(0), T(0), T(0)}, zhd[3] = {T(0), T(0), T(0)};
    #pragma unroll
                    for (int ic = 0; ic < KS; ++ic) {
                        const int li = row + ic;
                        T v[3], h[3];
    #pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            v[q] = sv[q * a.cap + li];
                            h[q] = sv[(3 + q) * a.cap + li];
                        }
                        T sn = m * sv[6 * a.cap + li];
    #pragma unroll
                        for (int q = 0; q < 3; ++q)
                            sn += (g01[q] + Gt.a[3 * q + 2] * d2[ic]) * h[q] + (c01[q] + Gc.a[3 * q + 2] * d2[ic]) * v[q];
                        zs += S.w[2][ic] * sn;
                        zsd += S.dw[2][ic] * sn;
    #pragma unroll
                        for (int q = 0; q < 3; ++q) {
                            zh[q] += S.w[2][ic] * h[q];
                            zhd[q] += w2d[ic] * h[q];
                        }
                    }
                    const T w0 = S.w[0][ia], w1 = S.w[1][ib], w01 = w0 * w1;
                    dxa.a[0] += invdx * S.dw[0][ia] * w1 * zs;
                    dxa.a[1] += invdx * w0 * S.dw[1][ib] * zs;
                    dxa.a[2] += invdx * w01 * zsd;
                    const T w01d0 = w01 * d0, w01d1 = w01 * d1;
    #pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        sh.a[q] += w01 * zh[q];
                        gA.a[3 * q] += w01d0 * zh[q];
                        gA.a[3 * q + 1] += w01d1 * zh[q];
                        gA.a[3 * q + 2] += w01 * zhd[q];
                    }
                }
            }
        } else {
    #pragma unroll
            for (int ia = 0; ia < KS; ++ia)
    #pragma unroll
                for (int ib = 0; ib < KS; ++ib)
    #pragma unroll
                    for (int ic = 0; ic < KS; ++ic) {
                        const int gi = S.b[0] + ia, gj = S.b[1] + ib, gk = S.b[2] + ic;
                        if (!in_grid(a, gi, gj, gk)) continue;
                        T v[3], h[3], gmass;
                        node_adjoint<T, REF>(a, gindex(a, gi, gj, gk), gi, gj, gk, v, h, gmass);
                        add(ia, ib, ic, v, h, gmass);
                    }
        }
        // F + dFc is loaded again here rather than held through the loop, and
        // the parked phase-1 terms are read back.  The index goes through an
        // empty asm statement so that the compiler reloads instead of keeping
        // the earlier values.
        int j = i;
        asm volatile("" : "+v"(j));
        m3<T> Ft2 = load_m3(a.st, a.ld, PF, j);
        if (a.ctl) Ft2 = madd(Ft2, load_m3(a.ctl, a.ld, 0, j));
        m3<T> Q;
        v3<T> gxp;
        {
            const T* park = reinterpret_cast<const T*>(reinterpret_cast<const char*>(smem_d) + a.park);
            const int pt = threadIdx.x, pb = blockDim.x;
#pragma unroll
            for (int q = 0; q < 9; ++q) Q.a[q] = park[q * pb + pt];
#pragma unroll
            for (int q = 0; q < 3; ++q) gxp.a[q] = park[(9 + q) * pb + pt];
        }
        // dL/dx_p: sum s_i grad w - Gt^T sum w h - D^{-1} gCraw^T v' + dL/dx'
        const v3<T> Gsh = mtv(Gt, sh);
        v3<T> gx;
    #pragma unroll
        for (int q = 0; q < 3; ++q) gx.a[q] = dxa.a[q] - Gsh.a[q] + gxp.a[q];
        const T dragf = m * (T(1) - a.dt * a.drag);
        const v3<T> gv = {{dragf * sh.a[0], dragf * sh.a[1], dragf * sh.a[2]}};
        // A_p = -k0 (P Ft^T + zeta tr C I) + m C: the stress term's gradient
        // in Ft is that of <-k0 gA, P(Ft) Ft^T>
        m3<T> gS;
        if constexpr (park_polar<T, MODEL>()) {
            const T* park = reinterpret_cast<const T*>(reinterpret_cast<const char*>(smem_d) + a.park);
            const int pt = threadIdx.x, pb = blockDim.x;
            m3<T> R, S;
#pragma unroll
            for (int q = 0; q < 9; ++q) R.a[q] = park[(12 + q) * pb + pt];
#pragma unroll
            for (int q = 0; q < 6; ++q) S.a[kSymUpper[q]] = park[(21 + q) * pb + pt];
            S.a[3] = S.a[1];  // (the stress gradient reads the upper triangle only)
            S.a[6] = S.a[2];
            S.a[7] = S.a[5];
            gS = Model<0>::kirchhoff_vjp(Model<0>::parts(Ft2, R, S), Ft2, lam, mu, mscale(-k0, gA));
        } else {
            gS = Model<MODEL>::kirchhoff_vjp(Ft2, lam, mu, mscale(-k0, gA));
        }
        m3<T> gC = mscale(m, gA);
        if (REF) {
            // bulk-viscosity adjoint on the diagonal entries only (:117-123)
            const T factor = -k0 * a.zeta;
            gC.a[0] += factor * gA.a[0];
            gC.a[4] += factor * gA.a[4];
            gC.a[8] += factor * gA.a[8];
        } else {
            const T tb = -k0 * a.zeta * mtrace(gA);
            gC.a[0] += tb; gC.a[4] += tb; gC.a[8] += tb;
        }
        const m3<T> gF = madd(Q, gS);
        const int io = !FUSED && a.opos ? a.opos[i] : i;
        store_v3(a.adj0, a.ld, PX, io, gx);
        store_v3(a.adj0, a.ld, PV, io, gv);
        store_m3(a.adj0, a.ld, PF, io, gF);
        store_m3(a.adj0, a.ld, PC, io, gC);
        if (FUSED) {
            cyt.gx = gx;
            cyt.gv = gv;
            cyt.gF = gF;
            cyt.gC = gC;
        }
        if (a.gctl) store_m3(a.gctl, a.ld, 0, i, madd(load_m3(a.gctl, a.ld, 0, j), gS));
    }
    if constexpr (FUSED) {
        v3<T> x1 = vzero<T>();
        m3<T> Ft1 = mzero<T>(), C1 = mzero<T>();
        if (valid) {
            x1 = load_v3(b->st, b->ld, PX, i);
            Ft1 = load_m3(b->st, b->ld, PF, i);
            if (b->ctl) Ft1 = madd(Ft1, load_m3(b->ctl, b->ld, 0, i));
            C1 = load_m3(a.st, a.ld, PC, i);
        }
        // The scatter window's reduction barrier (inside block_window) also
        // ends every wave's reads of the node values staged above, which its
        // accumulators overwrite; a barrier here is needed only when nothing
        // was staged, to keep the two reductions' tables apart.
        if (!win.lds) __syncthreads();
        adj_scatter_core<T, ORDER, REF>(*b, valid, x1, Ft1, C1, cyt, true, nullptr, reinterpret_cast<Acc*>(smem_d));
    }
}

template <typename T, int ORDER, int MODEL, bool REF>
__global__ void __launch_bounds__(256) k_adj_gather(KP<T> a) {
    extern __shared__ double smem_d[];
    adj_gather_block<T, ORDER, MODEL, REF, false>(a, nullptr, logical_block(), smem_d);
}

// Minimum waves per SIMD the compiler must fit the fused adjoint into: four
// (at most 128 VGPRs) for FP32 with the quadratic B-spline and Neo-Hookean
// stress, which needs 133 on its own and spills five values to scratch at
// 128 (measured at 1M particles: 158.3 -> 153.5 us per launch); no bound
// for the other forms (FP64 and the cubic / fixed-corotated ones hold many
// more registers, and would spill heavily).
template <typename T, int ORDER, int MODEL>
constexpr int adj_fused_waves() {
    return sizeof(T) == 4 && ORDER == 2 && MODEL == 1 ? 4 : 1;
}

// Adjoint of step t (gather) and of step t-1 (scatter) in one launch.
template <typename T, int ORDER, int MODEL, bool REF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(adj_fused_waves<T, ORDER, MODEL>())))
k_adj_fused(KP<T> a, KP<T> b) {
    extern __shared__ double smem_d[];
    const int lb = logical_block();
    adj_gather_block<T, ORDER, MODEL, REF, true>(a, &b, lb, smem_d);
    clear_share(a, lb);
}

template <typename T, int ORDER, int MODEL, bool REF>
__global__ void __launch_bounds__(256) k_adj_gather_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[inst];
    adj_gather_block<T, ORDER, MODEL, REF, false>(a, nullptr, lblock, smem_d);
}

// Batched k_adj_fused: tab holds two parameter blocks per instance (step t,
// then step t-1).
template <typename T, int ORDER, int MODEL, bool REF>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(adj_fused_waves<T, ORDER, MODEL>())))
k_adj_fused_batch(const KP<T>* tab, const int* bstart, int count) {
    extern __shared__ double smem_d[];
    int inst, lblock;
    if (!batch_block(bstart, count, inst, lblock)) return;
    const KP<T> a = tab[2 * inst];
    const KP<T> b = tab[2 * inst + 1];
    adj_gather_block<T, ORDER, MODEL, REF, true>(a, &b, lblock, smem_d);
    clear_share(a, lblock);
}

#if DMPM_IN(0)
// ---------------------------------------------------------------------------
// Deterministic two-pass scatter.  Pass 1 computes each particle's
// contribution terms and its base cell; a stable radix sort orders particles
// by (base cell, particle index); pass 2 runs one thread per grid node, which
// sums the contributions of the particles in the K^3 base cells that reach
// it, in that fixed order, and stores the node (no atomics, no clear).  The
// result is bitwise reproducible from run to run.

// Base-cell key with a 4-cell margin on each side; particles beyond it reach
// no grid node and are never gathered.
template <typename T>
DMPM_DI unsigned det_key(const KP<T>& a, const int* b) {
    int c[3];
    const int dm[3] = {a.nx, a.ny, a.nz};
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = min(max(b[q] + 4, 0), dm[q] + 7);
    return (unsigned)(((long long)c[0] * (a.ny + 8) + c[1]) * (a.nz + 8) + c[2]);
}

// Pass 1 of P2G: payload m (1) | mv (3) | A (9) per particle.
template <typename T, int ORDER, int MODEL, bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_det_prep_p2g(KP<T> a, unsigned* keys, int* idx, T* pay, long long pld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const v3<T> x = load_v3(a.st, a.ld, PX, i);
    Sten<T, ORDER> S;
    S.init(a, x);
    keys[i] = det_key(a, S.b);
    idx[i] = i;
    const T m = load_s(a.stat, a.ld, SM, i);
    pay[i] = m;
    if (!MASS_ONLY) {
        m3<T> A;
        v3<T> mv;
        m3<T> F = load_m3(a.st, a.ld, PF, i);
        if (a.ctl) F = madd(F, load_m3(a.ctl, a.ld, 0, i));
        p2g_terms<T, ORDER, MODEL>(a, i, load_v3(a.st, a.ld, PV, i), F, load_m3(a.st, a.ld, PC, i), m, A, mv);
#pragma unroll
        for (int q = 0; q < 3; ++q) pay[(1 + q) * pld + i] = mv.a[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) pay[(4 + q) * pld + i] = A.a[q];
    }
}

__global__ void k_det_cell_ranges(const unsigned* keys, int n, int* cstart, int* cend) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const unsigned k = keys[s];
    if (s == 0 || keys[s - 1] != k) cstart[k] = s;
    if (s == n - 1 || keys[s + 1] != k) cend[k] = s + 1;
}

// B-spline value/argument for the node gather, from |u| (any piece).
template <typename T, int ORDER>
DMPM_DI T det_N(T u) {
    const T ax = fabs(u);
    if (ORDER == 3) {
        if (ax < T(1)) return T(0.5) * ax * ax * ax - ax * ax + T(2) / T(3);
        if (ax < T(2)) return (T(2) - ax) * (T(2) - ax) * (T(2) - ax) / T(6);
        return T(0);
    }
    if (ax < T(0.5)) return T(0.75) - ax * ax;
    if (ax < T(1.5)) return T(0.5) * (T(1.5) - ax) * (T(1.5) - ax);
    return T(0);
}

// Visit, in fixed order, the particles whose stencil reaches node (i,j,k):
// fn(p, u0, u1, u2, w) with u = node - fx_p in cell units.
template <typename T, int ORDER, typename Fn>
DMPM_DI void det_visit(const KP<T>& a, int i, int j, int k, const int* cstart, const int* cend, const int* idx,
                       Fn&& fn) {
    constexpr int KS = Spline<ORDER>::K;
    const T inv = T(1) / a.dx;
    for (int da = 0; da < KS; ++da)
        for (int db = 0; db < KS; ++db)
            for (int dc = 0; dc < KS; ++dc) {
                const int b[3] = {i - KS + 1 + da, j - KS + 1 + db, k - KS + 1 + dc};
                if (b[0] + 4 < 0 || b[1] + 4 < 0 || b[2] + 4 < 0 || b[0] + 4 > a.nx + 7 || b[1] + 4 > a.ny + 7 ||
                    b[2] + 4 > a.nz + 7)
                    continue;
                const unsigned key = det_key(a, b);
                const int e = cend[key];
                for (int s = cstart[key]; s < e; ++s) {
                    const int p = idx[s];
                    const T u0 = (T)i - (a.st[PX * a.ld + p] - a.mn0) * inv;
                    const T u1 = (T)j - (a.st[(PX + 1) * a.ld + p] - a.mn1) * inv;
                    const T u2 = (T)k - (a.st[(PX + 2) * a.ld + p] - a.mn2) * inv;
                    const T w = det_N<T, ORDER>(u0) * det_N<T, ORDER>(u1) * det_N<T, ORDER>(u2);
                    fn(p, u0, u1, u2, w);
                }
            }
}

// Pass 2 of P2G: one thread per node.
template <typename T, int ORDER, bool MASS_ONLY>
__global__ void __launch_bounds__(256) k_det_gather_p2g(KP<T> a, const int* cstart, const int* cend, const int* idx,
                                                        const T* pay, long long pld) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long G = (long long)a.nx * a.ny * a.nz;
    if (g >= G) return;
    const int k = (int)(g % a.nz), j = (int)((g / a.nz) % a.ny), i = (int)(g / ((long long)a.ny * a.nz));
    T m = T(0), p0 = T(0), p1 = T(0), p2 = T(0);
    det_visit<T, ORDER>(a, i, j, k, cstart, cend, idx, [&](int q, T u0, T u1, T u2, T w) {
        m += w * pay[q];
        if (!MASS_ONLY) {
            const T d0 = u0 * a.dx, d1 = u1 * a.dx, d2 = u2 * a.dx;
            const T* A = pay + 4 * pld + q;
            p0 += w * (pay[pld + q] + A[0] * d0 + A[pld] * d1 + A[2 * pld] * d2);
            p1 += w * (pay[2 * pld + q] + A[3 * pld] * d0 + A[4 * pld] * d1 + A[5 * pld] * d2);
            p2 += w * (pay[3 * pld + q] + A[6 * pld] * d0 + A[7 * pld] * d1 + A[8 * pld] * d2);
        }
    });
    if (MASS_ONLY) {
        a.gm[g] = m;  // a mass-only grid: one value per node
    } else {
        T* n = a.gm + kNode * g;
        n[0] = m;
        n[1] = p0;
        n[2] = p1;
        n[3] = p2;
    }
}

// Pass 1 of the adjoint scatter: payload gvtot (3) | D^{-1} gCraw (9).
template <typename T, int ORDER, bool REF>
__global__ void __launch_bounds__(256) k_det_prep_adj(KP<T> a, unsigned* keys, int* idx, T* pay, long long pld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const v3<T> x = load_v3(a.st, a.ld, PX, i);
    Sten<T, ORDER> S;
    S.init(a, x);
    keys[i] = det_key(a, S.b);
    idx[i] = i;
    Win nowin;
    nowin.lds = false;
    v3<T> vn;
    m3<T> Craw;
    g2p_gather<T, ORDER>(a, nowin, nullptr, S, vn, Craw);
    m3<T> Ft = load_m3(a.st, a.ld, PF, i);
    if (a.ctl) Ft = madd(Ft, load_m3(a.ctl, a.ld, 0, i));
    const Trust<T> tr = trust_region(a, Ft, Craw);
    const Phase1<T, REF> p1 = phase1<T, REF>(a, i, Ft, Craw, tr);
    const T dinv = Spline<ORDER>::dinv(a.dx);
#pragma unroll
    for (int q = 0; q < 3; ++q) pay[q * pld + i] = p1.gvtot.a[q];
#pragma unroll
    for (int q = 0; q < 9; ++q) pay[(3 + q) * pld + i] = dinv * p1.gCraw.a[q];
}

template <typename T, int ORDER>
__global__ void __launch_bounds__(256) k_det_gather_adj(KP<T> a, const int* cstart, const int* cend, const int* idx,
                                                        const T* pay, long long pld) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long G = (long long)a.nx * a.ny * a.nz;
    if (g >= G) return;
    const int k = (int)(g % a.nz), j = (int)((g / a.nz) % a.ny), i = (int)(g / ((long long)a.ny * a.nz));
    T g0 = T(0), g1 = T(0), g2 = T(0);
    det_visit<T, ORDER>(a, i, j, k, cstart, cend, idx, [&](int q, T u0, T u1, T u2, T w) {
        const T d0 = u0 * a.dx, d1 = u1 * a.dx, d2 = u2 * a.dx;
        const T* Gc = pay + 3 * pld + q;
        g0 += w * (pay[q] + Gc[0] * d0 + Gc[pld] * d1 + Gc[2 * pld] * d2);
        g1 += w * (pay[pld + q] + Gc[3 * pld] * d0 + Gc[4 * pld] * d1 + Gc[5 * pld] * d2);
        g2 += w * (pay[2 * pld + q] + Gc[6 * pld] * d0 + Gc[7 * pld] * d1 + Gc[8 * pld] * d2);
    });
    T* n = a.ggv + kNode * g;
    n[0] = g0;
    n[1] = g1;
    n[2] = g2;
    n[3] = T(0);
}

// ---------------------------------------------------------------------------
// Small element-wise kernels.
template <typename T>
__global__ void k_fill_zero(T* p, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = T(0);
}

// Zero n16 16-byte words (grids: 16-byte aligned, sizes a multiple of
// 64 elements) with one dwordx4 store per word.
__global__ void k_fill_zero16(uint4* p, long long n16) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (; i < n16; i += stride) p[i] = z;
}

template <typename T, typename S>
__global__ void k_import(const S* src, T* dst, const int* perm, int n, int nc, long long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long s = (long long)(perm ? perm[i] : i) * nc;
    for (int c = 0; c < nc; ++c) dst[c * ld + i] = (T)src[s + c];
}

template <typename T, typename S>
__global__ void k_accumulate(const S* src, T* dst, const int* perm, int n, int nc, long long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long s = (long long)(perm ? perm[i] : i) * nc;
    for (int c = 0; c < nc; ++c) dst[c * ld + i] += (T)src[s + c];
}

template <typename T, typename D>
__global__ void k_export(const T* src, D* dst, const int* perm, int n, int nc, long long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long d = (long long)(perm ? perm[i] : i) * nc;
    for (int c = 0; c < nc; ++c) dst[d + c] = (D)src[c * ld + i];
}

template <typename T, typename D>
__global__ void k_export_Ftot(const T* st, const T* ctl, D* dst, const int* perm, int n, long long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long d = (long long)(perm ? perm[i] : i) * 9;
    for (int c = 0; c < 9; ++c) dst[d + c] = (D)(st[(PF + c) * ld + i] + (ctl ? ctl[c * ld + i] : T(0)));
}

// First Piola-Kirchhoff stress of F + dFc under MODEL (the reference keeps it
// as MaterialPoint::P after P_op_1, ForwardSimulation.cpp:16-24), exported in
// caller order.
template <typename T, typename D, int MODEL>
__global__ void k_export_stress(const T* st, const T* ctl, const T* stat, D* dst, const int* perm, int n,
                                long long ld) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    m3<T> F = load_m3(st, ld, PF, i);
    if (ctl) F = madd(F, load_m3(ctl, ld, 0, i));
    const m3<T> P = Model<MODEL>::pk1(F, stat[SLAM * ld + i], stat[SMU * ld + i]);
    const long long d = (long long)(perm ? perm[i] : i) * 9;
    for (int c = 0; c < 9; ++c) dst[d + c] = (D)P.a[c];
}

// Generalised winding number of points with respect to a triangle mesh (the
// inside test of the mesh sampler, GeometryLoading.cpp:61-77):
// w(p) = (1/4pi) sum_t Omega_t(p), the signed solid angle of triangle
// (A, B, C) seen from p following A. Van Oosterom and J. Strackee, "The solid
// angle of a plane triangle", IEEE Trans. Biomed. Eng. 30(2), 1983:
// tan(Omega/2) = a.(b x c) / (|a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|),
// a = A - p, b = B - p, c = C - p.  Block (x, y) sums the triangles
// [y per, (y+1) per) for 256 points, staging WIND_TILE triangles at a time in
// LDS (every lane reads the same triangle, a broadcast); partial[y np + i]
// holds its sum and k_winding_sum adds the chunks in a fixed order.
constexpr int WIND_TILE = 256;
__global__ void __launch_bounds__(256) k_winding(const double* tri, int nf, int per, const double* P, int np,
                                                 double* partial) {
    __shared__ double st[WIND_TILE * 9];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < np;
    double px = 0, py = 0, pz = 0;
    if (valid) {
        px = P[3 * (long long)i];
        py = P[3 * (long long)i + 1];
        pz = P[3 * (long long)i + 2];
    }
    const int t0 = blockIdx.y * per, t1 = min(nf, t0 + per);
    double acc = 0;
    for (int base = t0; base < t1; base += WIND_TILE) {
        const int cnt = min(WIND_TILE, t1 - base);
        __syncthreads();
        for (int q = threadIdx.x; q < cnt * 9; q += blockDim.x) st[q] = tri[(long long)base * 9 + q];
        __syncthreads();
        if (!valid) continue;
        for (int k = 0; k < cnt; ++k) {
            const double* T = st + 9 * k;
            const double ax = T[0] - px, ay = T[1] - py, az = T[2] - pz;
            const double bx = T[3] - px, by = T[4] - py, bz = T[5] - pz;
            const double cx = T[6] - px, cy = T[7] - py, cz = T[8] - pz;
            const double la = sqrt(ax * ax + ay * ay + az * az);
            const double lb = sqrt(bx * bx + by * by + bz * bz);
            const double lc = sqrt(cx * cx + cy * cy + cz * cz);
            const double det = ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
            const double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (bx * cx + by * cy + bz * cz) * la +
                               (cx * ax + cy * ay + cz * az) * lb;
            acc += 2.0 * atan2(det, den);
        }
    }
    if (valid) partial[(long long)blockIdx.y * np + i] = acc;
}

__global__ void k_winding_sum(const double* partial, int chunks, int np, double* w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    double s = 0;
    for (int c = 0; c < chunks; ++c) s += partial[(long long)c * np + i];
    w[i] = s / (4.0 * 3.14159265358979323846);
}

template <typename T>
__global__ void k_copy(T* dst, const T* src, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

// Block sum of doubles; result valid in thread 0.
DMPM_DI double block_sum(double v) {
    __shared__ double red[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[q];
    }
    return s;
}

template <typename T>
__global__ void k_mass_sums(const T* gm, const T* tm, long long G, double* partial) {
    double c = 0, t = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += (long long)gridDim.x * blockDim.x) {
        c += (double)gm[i];
        t += (double)tm[i];
    }
    c = block_sum(c);
    t = block_sum(t);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = c;
        partial[2 * blockIdx.x + 1] = t;
    }
}

// EndLayerMassLoss node terms (CompGraph.cpp:60-93).
template <typename T>
__global__ void k_mass_loss_nodes(const T* gm, const T* tm, long long G, const double* sums, T* dldm,
                                  double* partial) {
    const T sum_c = (T)sums[0], sum_t = (T)sums[1];
    const T s = (sum_t + T(1e-6)) / (sum_c + T(1e-6));
    const T eps = T(1e-4), min_mass = T(1e-3);
    double acc = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < G; i += (long long)gridDim.x * blockDim.x) {
        const T cms = s * gm[i];
        const T ld = log(cms + T(1) + eps) - log(tm[i] + T(1) + eps);
        acc += (double)(T(0.5) * ld * ld);
        T d = (s / (cms + T(1) + eps)) * ld;
        if (cms < min_mass) {
            const T diff = min_mass - cms;
            acc += (double)(diff * diff);
            d += T(-2) * diff * s;
        }
        dldm[i] = d;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// Per-particle dL/dx of EndLayerMassLoss (CompGraph.cpp:95-125); the other
// adjoint fields of the last layer are zeroed (:122-124).
template <typename T, int ORDER>
__global__ void k_mass_loss_particles(KP<T> a, const T* dldm, const T* tm) {
    constexpr int KS = Spline<ORDER>::K;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const v3<T> x = load_v3(a.st, a.ld, PX, i);
    Sten<T, ORDER> S;
    S.init(a, x);
    const T m = load_s(a.stat, a.ld, SM, i);
    v3<T> g = vzero<T>();
    const T invdx = T(-1) / a.dx;
    for (int ia = 0; ia < KS; ++ia)
        for (int ib = 0; ib < KS; ++ib)
            for (int ic = 0; ic < KS; ++ic) {
                const int gi = S.b[0] + ia, gj = S.b[1] + ib, gk = S.b[2] + ic;
                if (!in_grid(a, gi, gj, gk)) continue;
                const long long gg = gindex(a, gi, gj, gk);
                const T pen = tm[gg] > T(1e-12) ? T(1) : T(5);
                const T c = pen * m * dldm[gg];
                g.a[0] += c * invdx * S.dw[0][ia] * S.w[1][ib] * S.w[2][ic];
                g.a[1] += c * invdx * S.w[0][ia] * S.dw[1][ib] * S.w[2][ic];
                g.a[2] += c * invdx * S.w[0][ia] * S.w[1][ib] * S.dw[2][ic];
            }
    store_v3(a.adj0, a.ld, PX, i, g);
    for (int q = 3; q < NSTATE; ++q) a.adj0[q * a.ld + i] = T(0);
}

// CalculatePointCloudVolumes particle pass (ForwardSimulation.cpp:241-269).
template <typename T, int ORDER>
__global__ void k_volumes(KP<T> a, T* vol) {
    constexpr int KS = Spline<ORDER>::K;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const v3<T> x = load_v3(a.st, a.ld, PX, i);
    Sten<T, ORDER> S;
    S.init(a, x);
    T mg = T(0);
    for (int ia = 0; ia < KS; ++ia)
        for (int ib = 0; ib < KS; ++ib)
            for (int ic = 0; ic < KS; ++ic) {
                const int gi = S.b[0] + ia, gj = S.b[1] + ib, gk = S.b[2] + ic;
                if (!in_grid(a, gi, gj, gk)) continue;
                mg += S.w[0][ia] * S.w[1][ib] * S.w[2][ic] * a.gm[gindex(a, gi, gj, gk)];
            }
    const T cell = a.dx * a.dx * a.dx;
    const T rho = mg / fmax(cell, T(1e-12));
    const T eps = sizeof(T) == 4 ? T(1.1920929e-07) : T(2.220446049250313e-16);
    vol[i] = rho > eps ? load_s(a.stat, a.ld, SM, i) / rho : T(0);
}

template <typename T>
__global__ void k_sqnorm(const T* p, int n, int nc, long long ld, double* partial) {
    double acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int q = 0; q < nc; ++q) {
            const double v = (double)p[q * ld + i];
            acc += v * v;
        }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

__global__ void k_reduce_partials(const double* partial, int nblocks, int stride, double* out) {
    // one thread per output component, fixed order: deterministic
    const int c = threadIdx.x;
    if (c >= stride) return;
    double s = 0;
    for (int b = 0; b < nblocks; ++b) s += partial[b * stride + c];
    out[c] = s;
}

// Descend_Adam (PointCloud.cpp:83-147).
template <typename T>
__global__ void k_adam(const T* dLdF, T* dFc, T* mom, T* vec, T* vmax, int n, long long ld, const double* sqnorm,
                       T alpha, T b1, T b2, T eps, int timestep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T gn = (T)sqrt(sqnorm[0]);
    const T inv_gn = T(1) / fmax(fmax(gn, T(1e-12)), T(1e-12));
    const int t = timestep < 1 ? 1 : timestep;
    const T c1 = T(1) - (T)pow((double)b1, t);
    const T c2 = T(1) - (T)pow((double)b2, t);
    T step[9], xo[9];
    T sn = T(0), xn = T(0);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const long long o = q * ld + i;
        const T g = dLdF[o] * inv_gn;
        const T mq = b1 * mom[o] + (T(1) - b1) * g;
        const T vq = b2 * vec[o] + (T(1) - b2) * g * g;
        mom[o] = mq;
        vec[o] = vq;
        const T vm = fmax(vmax[o], vq / c2);
        vmax[o] = vm;
        const T ds = fmax(sqrt(vm), T(1e-8));
        step[q] = (mq / c1) / (ds + eps);
        sn += step[q] * step[q];
        xo[q] = dFc[o];
        xn += xo[q] * xo[q];
    }
    sn = sqrt(sn);
    xn = sqrt(xn);
    const T cap = fmax(T(5e-3), T(5e-2) * fmax(T(1), xn));
    T f = T(1);
    if (sn > T(0)) {
        const T scaled = alpha * sn;
        if (scaled > cap) f = cap / scaled;
    }
    const T small_cap = fmax(T(1e-4), T(1e-3) * fmax(T(1), xn));
    const bool near_stationary = alpha * sn <= small_cap;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        T xq = xo[q] - alpha * (f * step[q]);
        if (near_stationary) xq = xo[q] + T(0.2) * (xq - xo[q]);
        dFc[q * ld + i] = xq;
    }
}

#endif  // DMPM_IN(0)

// ---------------------------------------------------------------------------
// Launchers.

// LDS of the adjoint gather: its window (budget bytes), then 12 values per
// thread parked through the node loop.  With FP32 the window budget is held
// to 27 KiB so that four 256-thread blocks (window + 12 KiB parked) share a
// CU's 160 KiB.

template <typename T>
static int adj_window_budget(int lds_bytes, int model) {
    // FP32: the window plus the parked values fit four blocks per CU (three
    // for the fixed-corotated model, which parks 27 values and holds three
    // waves per SIMD)
    return sizeof(T) == 4 ? std::min(lds_bytes, (model == 0 ? 24 : 27) * 1024) : lds_bytes;
}
template <typename T>
static size_t adj_gather_lds(int budget, int block, int model) {
    return (size_t)((budget + 15) & ~15) + (size_t)park_values<T>(model) * block * sizeof(T);
}
template <typename T>
static KP<T> adj_gather_kp(const Launch& L) {
    Launch Lw = L;
    Lw.lds_bytes = adj_window_budget<T>(L.lds_bytes, L.ph.model);
    KP<T> k = make_kp<T>(Lw, 7 * (int)sizeof(T));
    k.park = (Lw.lds_bytes + 15) & ~15;
    return k;
}

#if DMPM_IN(1) || DMPM_IN(2)
template <typename T>
static void launch_p2g(const Launch& L, bool mass_only, hipStream_t s) {
    const int na = mass_only ? 1 : 4;
    KP<T> k = make_kp<T>(L, na * (int)sizeof(Acc));
    const unsigned nb = launch_blocks(L);
    const size_t sh = (size_t)k.cap * na * sizeof(Acc);
#define P2G_CASE(O, M, MO) hipLaunchKernelGGL((k_p2g<T, O, M, MO>), dim3(nb), dim3(L.block), sh, s, k)
    const int o = L.ph.order, md = L.ph.model;
    if (mass_only) {
        if (o == 3) P2G_CASE(3, 0, true); else P2G_CASE(2, 0, true);
    } else if (o == 3) {
        if (md == 0) P2G_CASE(3, 0, false); else P2G_CASE(3, 1, false);
    } else {
        if (md == 0) P2G_CASE(2, 0, false); else P2G_CASE(2, 1, false);
    }
#undef P2G_CASE
}

template <typename T>
static void launch_g2p(const Launch& L, hipStream_t s) {
    KP<T> k = make_kp<T>(L, 3 * (int)sizeof(T));
    const unsigned nb = launch_blocks(L);
    const size_t sh = (size_t)k.cap * 3 * sizeof(T);
    if (L.ph.order == 3) hipLaunchKernelGGL((k_g2p<T, 3>), dim3(nb), dim3(L.block), sh, s, k);
    else hipLaunchKernelGGL((k_g2p<T, 2>), dim3(nb), dim3(L.block), sh, s, k);
}

template <typename T>
static void launch_step(const Launch& L, hipStream_t s) {
    // staging needs 3 T per node, the P2G accumulators 4 Acc per node
    KP<T> k = make_kp<T>(L, 4 * (int)sizeof(Acc));
    const unsigned nb = launch_blocks(L);
    const size_t sh = (size_t)k.cap * 4 * sizeof(Acc);
    const int o = L.ph.order, md = L.ph.model;
#define ST_CASE(O, M) hipLaunchKernelGGL((k_step<T, O, M>), dim3(nb), dim3(L.block), sh, s, k)
    if (o == 3) { if (md == 0) ST_CASE(3, 0); else ST_CASE(3, 1); }
    else { if (md == 0) ST_CASE(2, 0); else ST_CASE(2, 1); }
#undef ST_CASE
}

template <typename T>
void launch_fwd(int kind, const Launch& L, hipStream_t s) {
    switch (kind) {
        case OP_P2G: launch_p2g<T>(L, false, s); break;
        case OP_P2G_MASS: launch_p2g<T>(L, true, s); break;
        case OP_G2P: launch_g2p<T>(L, s); break;
        case OP_STEP: launch_step<T>(L, s); break;
        default: break;
    }
}

template <typename T>
void batch_fwd(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab_, const int* bstart,
               int count, hipStream_t s) {
    const KP<T>* tab = (const KP<T>*)tab_;
    const int o = ph.order, md = ph.model;
#define BATCH(KERNEL) hipLaunchKernelGGL(KERNEL, dim3(nb), dim3(block), sh, s, tab, bstart, count)
    switch (kind) {
        case OP_P2G:
            if (o == 3) { if (md == 0) BATCH((k_p2g_batch<T, 3, 0, false>)); else BATCH((k_p2g_batch<T, 3, 1, false>)); }
            else { if (md == 0) BATCH((k_p2g_batch<T, 2, 0, false>)); else BATCH((k_p2g_batch<T, 2, 1, false>)); }
            break;
        case OP_P2G_MASS:
            if (o == 3) BATCH((k_p2g_batch<T, 3, 0, true>)); else BATCH((k_p2g_batch<T, 2, 0, true>));
            break;
        case OP_G2P:
            if (o == 3) BATCH((k_g2p_batch<T, 3>)); else BATCH((k_g2p_batch<T, 2>));
            break;
        case OP_STEP:
            if (o == 3) { if (md == 0) BATCH((k_step_batch<T, 3, 0>)); else BATCH((k_step_batch<T, 3, 1>)); }
            else { if (md == 0) BATCH((k_step_batch<T, 2, 0>)); else BATCH((k_step_batch<T, 2, 1>)); }
            break;
        default: break;
    }
#undef BATCH
}
#endif
#if DMPM_IN(1)
template void launch_fwd<float>(int, const Launch&, hipStream_t);
template void batch_fwd<float>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif
#if DMPM_IN(2)
template void launch_fwd<double>(int, const Launch&, hipStream_t);
template void batch_fwd<double>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif

#if DMPM_IN(3) || DMPM_IN(4)
template <typename T>
static void launch_adj_scatter(const Launch& L, hipStream_t s) {
    const bool ref = L.ph.adjoint == 1;
    // the kernel reads layer t+1 (reads_next): its window holds the 3
    // accumulators only, so it may span more nodes in the same LDS
    const bool nxt = L.nxt != nullptr && (ref || !L.ph.trust_region);
    const int per_node = nxt ? 3 * (int)sizeof(Acc) : 3 * (int)sizeof(T) + 3 * (int)sizeof(Acc) + 8;
    KP<T> k = make_kp<T>(L, per_node);
    const unsigned nb = launch_blocks(L);
    const size_t sh = nxt ? (size_t)k.cap * 3 * sizeof(Acc) : (size_t)k.cap * (3 * sizeof(T) + 3 * sizeof(Acc)) + 8;
#define AS_CASE(O, R) hipLaunchKernelGGL((k_adj_scatter<T, O, R>), dim3(nb), dim3(L.block), sh, s, k)
    if (L.ph.order == 3) { if (ref) AS_CASE(3, true); else AS_CASE(3, false); }
    else { if (ref) AS_CASE(2, true); else AS_CASE(2, false); }
#undef AS_CASE
}

template <typename T>
static void launch_adj_gather(const Launch& L, hipStream_t s) {
    KP<T> k = adj_gather_kp<T>(L);
    const unsigned nb = launch_blocks(L);
    const size_t sh = adj_gather_lds<T>(adj_window_budget<T>(L.lds_bytes, L.ph.model), L.block, L.ph.model);
    const bool ref = L.ph.adjoint == 1;
    const int o = L.ph.order, md = L.ph.model;
#define AG_CASE(O, M, R) hipLaunchKernelGGL((k_adj_gather<T, O, M, R>), dim3(nb), dim3(L.block), sh, s, k)
    if (o == 3) {
        if (md == 0) { if (ref) AG_CASE(3, 0, true); else AG_CASE(3, 0, false); }
        else { if (ref) AG_CASE(3, 1, true); else AG_CASE(3, 1, false); }
    } else {
        if (md == 0) { if (ref) AG_CASE(2, 0, true); else AG_CASE(2, 0, false); }
        else { if (ref) AG_CASE(2, 1, true); else AG_CASE(2, 1, false); }
    }
#undef AG_CASE
}

template <typename T>
void launch_adj(int kind, const Launch& L, hipStream_t s) {
    if (kind == OP_ADJ_SCATTER) launch_adj_scatter<T>(L, s);
    else if (kind == OP_ADJ_GATHER) launch_adj_gather<T>(L, s);
}

template <typename T>
void batch_adj(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab_, const int* bstart,
               int count, hipStream_t s) {
    const KP<T>* tab = (const KP<T>*)tab_;
    const int o = ph.order, md = ph.model;
    const bool ref = ph.adjoint == 1;
#define BATCH(KERNEL) hipLaunchKernelGGL(KERNEL, dim3(nb), dim3(block), sh, s, tab, bstart, count)
    switch (kind) {
        case OP_ADJ_SCATTER:
            if (o == 3) { if (ref) BATCH((k_adj_scatter_batch<T, 3, true>)); else BATCH((k_adj_scatter_batch<T, 3, false>)); }
            else { if (ref) BATCH((k_adj_scatter_batch<T, 2, true>)); else BATCH((k_adj_scatter_batch<T, 2, false>)); }
            break;
        case OP_ADJ_GATHER:
            if (o == 3) {
                if (md == 0) { if (ref) BATCH((k_adj_gather_batch<T, 3, 0, true>)); else BATCH((k_adj_gather_batch<T, 3, 0, false>)); }
                else { if (ref) BATCH((k_adj_gather_batch<T, 3, 1, true>)); else BATCH((k_adj_gather_batch<T, 3, 1, false>)); }
            } else {
                if (md == 0) { if (ref) BATCH((k_adj_gather_batch<T, 2, 0, true>)); else BATCH((k_adj_gather_batch<T, 2, 0, false>)); }
                else { if (ref) BATCH((k_adj_gather_batch<T, 2, 1, true>)); else BATCH((k_adj_gather_batch<T, 2, 1, false>)); }
            }
            break;
        default: break;
    }
#undef BATCH
}
#endif
#if DMPM_IN(3)
template void launch_adj<float>(int, const Launch&, hipStream_t);
template void batch_adj<float>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif
#if DMPM_IN(4)
template void launch_adj<double>(int, const Launch&, hipStream_t);
template void batch_adj<double>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif

#if DMPM_IN(5) || DMPM_IN(6)
template <typename T>
static void launch_adj_fused(const Launch& L, const Launch& L2, hipStream_t s) {
    KP<T> k = adj_gather_kp<T>(L);
    Launch Lw2 = L2;
    Lw2.lds_bytes = adj_window_budget<T>(L2.lds_bytes, L2.ph.model);
    KP<T> k2 = make_kp<T>(Lw2, 3 * (int)sizeof(Acc));
    const unsigned nb = launch_blocks(L);
    const size_t sh = std::max(adj_gather_lds<T>(adj_window_budget<T>(L.lds_bytes, L.ph.model), L.block, L.ph.model),
                               (size_t)k2.cap * 3 * sizeof(Acc));
    const bool ref = L.ph.adjoint == 1;
    const int o = L.ph.order, md = L.ph.model;
#define AF_CASE(O, M, R) hipLaunchKernelGGL((k_adj_fused<T, O, M, R>), dim3(nb), dim3(L.block), sh, s, k, k2)
    if (o == 3) {
        if (md == 0) { if (ref) AF_CASE(3, 0, true); else AF_CASE(3, 0, false); }
        else { if (ref) AF_CASE(3, 1, true); else AF_CASE(3, 1, false); }
    } else {
        if (md == 0) { if (ref) AF_CASE(2, 0, true); else AF_CASE(2, 0, false); }
        else { if (ref) AF_CASE(2, 1, true); else AF_CASE(2, 1, false); }
    }
#undef AF_CASE
}

template <typename T>
void launch_fused(const Launch& L, const Launch& L2, hipStream_t s) {
    launch_adj_fused<T>(L, L2, s);
}

template <typename T>
void batch_fused(int kind, const Phys& ph, int block, size_t sh, unsigned nb, const void* tab_, const int* bstart,
                 int count, hipStream_t s) {
    const KP<T>* tab = (const KP<T>*)tab_;
    const int o = ph.order, md = ph.model;
    const bool ref = ph.adjoint == 1;
    if (kind != OP_ADJ_FUSED) return;
#define BATCH(KERNEL) hipLaunchKernelGGL(KERNEL, dim3(nb), dim3(block), sh, s, tab, bstart, count)
    if (o == 3) {
        if (md == 0) { if (ref) BATCH((k_adj_fused_batch<T, 3, 0, true>)); else BATCH((k_adj_fused_batch<T, 3, 0, false>)); }
        else { if (ref) BATCH((k_adj_fused_batch<T, 3, 1, true>)); else BATCH((k_adj_fused_batch<T, 3, 1, false>)); }
    } else {
        if (md == 0) { if (ref) BATCH((k_adj_fused_batch<T, 2, 0, true>)); else BATCH((k_adj_fused_batch<T, 2, 0, false>)); }
        else { if (ref) BATCH((k_adj_fused_batch<T, 2, 1, true>)); else BATCH((k_adj_fused_batch<T, 2, 1, false>)); }
    }
#undef BATCH
}
#endif
#if DMPM_IN(5)
template void launch_fused<float>(const Launch&, const Launch&, hipStream_t);
template void batch_fused<float>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif
#if DMPM_IN(6)
template void launch_fused<double>(const Launch&, const Launch&, hipStream_t);
template void batch_fused<double>(int, const Phys&, int, size_t, unsigned, const void*, const int*, int, hipStream_t);
#endif

#if DMPM_IN(0)
// ---- batched launches --------------------------------------------------------
// Bytes of LDS per window node of each kernel op (the window capacity is the
// LDS budget divided by this), and the dynamic LDS a block of it takes.
template <typename T>
static int node_bytes(int kind) {
    switch (kind) {
        case OP_P2G: return 4 * (int)sizeof(Acc);
        case OP_P2G_MASS: return (int)sizeof(Acc);
        case OP_G2P: return 3 * (int)sizeof(T);
        case OP_STEP: return 4 * (int)sizeof(Acc);
        case OP_ADJ_SCATTER: return 3 * (int)sizeof(T) + 3 * (int)sizeof(Acc) + 8;
        case OP_ADJ_GATHER: return 7 * (int)sizeof(T);
        case OP_ADJ_FUSED: return 7 * (int)sizeof(T);  // the gather's window; the scatter's uses 3 Acc
        default: return 0;
    }
}
template <typename T>
static size_t block_lds(int kind, int cap) {
    if (kind == OP_ADJ_SCATTER) return (size_t)cap * (3 * sizeof(T) + 3 * sizeof(Acc)) + 8;
    return (size_t)cap * node_bytes<T>(kind);
}

size_t kp_bytes(Prec pr) { return pr == PREC_F32 ? sizeof(KP<float>) : sizeof(KP<double>); }

template <typename T>
static void build_kp_t(int kind, const Launch& L, const Launch* L2, void* dst) {
    const bool adj = kind == OP_ADJ_GATHER || kind == OP_ADJ_FUSED;
    KP<T> k = adj ? adj_gather_kp<T>(L) : make_kp<T>(L, node_bytes<T>(kind));
    k.zblk = instance_blocks(L);  // the instance's blocks of a batched launch
    if (k.zero && k.zblk <= 0) throw std::runtime_error("a clear over an instance without blocks");
    k.zper = k.zblk > 0 ? (k.zero16 + k.zblk - 1) / k.zblk : 0;
    std::memcpy(dst, &k, sizeof(k));
    if (kind == OP_ADJ_FUSED) {
        Launch Lw2 = *L2;
        Lw2.lds_bytes = adj_window_budget<T>(L2->lds_bytes, L2->ph.model);
        const KP<T> k2 = make_kp<T>(Lw2, 3 * (int)sizeof(Acc));
        std::memcpy((char*)dst + sizeof(k), &k2, sizeof(k2));
    }
}

void build_kp(Prec pr, int kind, const Launch& L, const Launch* L2, void* dst) {
    if (pr == PREC_F32) build_kp_t<float>(kind, L, L2, dst);
    else build_kp_t<double>(kind, L, L2, dst);
}

int instance_blocks(const Launch& L) {
    if (L.bst) return L.nblk;
    return L.n > 0 ? (L.n + L.block - 1) / L.block : 0;
}

template <typename T>
static void launch_batch_t(int kind, const Phys& ph, int lds_bytes, int block, const void* tab, const int* bstart,
                           int count, int total, hipStream_t s) {
    const unsigned nb = (unsigned)std::max(8, (total + 7) & ~7);
    const int cap = lds_bytes / node_bytes<T>(kind);
    size_t sh = block_lds<T>(kind, cap);
    if (kind == OP_ADJ_GATHER || kind == OP_ADJ_FUSED) {
        const int budget = adj_window_budget<T>(lds_bytes, ph.model);
        sh = adj_gather_lds<T>(budget, block, ph.model);
        if (kind == OP_ADJ_FUSED) sh = std::max(sh, (size_t)(budget / (3 * (int)sizeof(Acc))) * 3 * sizeof(Acc));
    }
    switch (kind) {
        case OP_P2G:
        case OP_P2G_MASS:
        case OP_G2P:
        case OP_STEP: batch_fwd<T>(kind, ph, block, sh, nb, tab, bstart, count, s); break;
        case OP_ADJ_SCATTER:
        case OP_ADJ_GATHER: batch_adj<T>(kind, ph, block, sh, nb, tab, bstart, count, s); break;
        case OP_ADJ_FUSED: batch_fused<T>(kind, ph, block, sh, nb, tab, bstart, count, s); break;
        default: break;
    }
}

void launch_batch(Prec pr, int kind, const Phys& ph, int lds_bytes, int block, const void* tab, const int* bstart,
                  int count, int total, hipStream_t s) {
    if (count <= 0 || total <= 0) return;
    if (pr == PREC_F32) launch_batch_t<float>(kind, ph, lds_bytes, block, tab, bstart, count, total, s);
    else launch_batch_t<double>(kind, ph, lds_bytes, block, tab, bstart, count, total, s);
}

// Segment kernels (blockIdx.y = range); zero and copy move 16-byte words.
__global__ void k_zero_segs(const Seg* segs, int esz) {
    const Seg g = segs[blockIdx.y];
    const long long n16 = g.n * esz / 16;
    uint4* d = (uint4*)g.dst;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
        d[i] = z;
}
__global__ void k_copy_segs(const Seg* segs, int esz) {
    const Seg g = segs[blockIdx.y];
    const long long n16 = g.n * esz / 16;
    uint4* d = (uint4*)g.dst;
    const uint4* src = (const uint4*)g.src;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
        d[i] = src[i];
}
template <typename T>
__global__ void k_accum_segs(const Seg* segs) {
    const Seg g = segs[blockIdx.y];
    T* d = (T*)g.dst;
    const T* src = (const T*)g.src;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < g.n; i += (long long)gridDim.x * blockDim.x)
        d[i] += src[i];
}

void launch_segs(Prec pr, int kind, const Seg* segs, int count, long long max_n, hipStream_t s) {
    if (count <= 0 || max_n <= 0) return;
    const int esz = pr == PREC_F32 ? 4 : 8;
    const long long words = kind == OP_ACCUM ? max_n : max_n * esz / 16;
    const unsigned gx = (unsigned)std::max(1LL, std::min((words + 1023) / 1024, 256LL));
    const dim3 grid(gx, (unsigned)count);
    if (kind == OP_ZERO) hipLaunchKernelGGL(k_zero_segs, grid, dim3(256), 0, s, segs, esz);
    else if (kind == OP_COPY) hipLaunchKernelGGL(k_copy_segs, grid, dim3(256), 0, s, segs, esz);
    else if (pr == PREC_F32) hipLaunchKernelGGL(k_accum_segs<float>, grid, dim3(256), 0, s, segs);
    else hipLaunchKernelGGL(k_accum_segs<double>, grid, dim3(256), 0, s, segs);
}

static inline unsigned ew_blocks(long long n, int bs = 256) {
    long long nb = (n + bs - 1) / bs;
    if (nb < 1) nb = 1;
    if (nb > 65536) nb = 65536;
    return (unsigned)nb;
}

void p2g(Prec pr, const Launch& L, bool mass_only, hipStream_t s) {
    const int kind = mass_only ? OP_P2G_MASS : OP_P2G;
    if (pr == PREC_F32) launch_fwd<float>(kind, L, s);
    else launch_fwd<double>(kind, L, s);
}
void g2p(Prec pr, const Launch& L, hipStream_t s) {
    if (pr == PREC_F32) launch_fwd<float>(OP_G2P, L, s);
    else launch_fwd<double>(OP_G2P, L, s);
}
void step(Prec pr, const Launch& L, hipStream_t s) {
    if (pr == PREC_F32) launch_fwd<float>(OP_STEP, L, s);
    else launch_fwd<double>(OP_STEP, L, s);
}
void adj_scatter(Prec pr, const Launch& L, hipStream_t s) {
    if (pr == PREC_F32) launch_adj<float>(OP_ADJ_SCATTER, L, s);
    else launch_adj<double>(OP_ADJ_SCATTER, L, s);
}
void adj_gather(Prec pr, const Launch& L, hipStream_t s) {
    if (pr == PREC_F32) launch_adj<float>(OP_ADJ_GATHER, L, s);
    else launch_adj<double>(OP_ADJ_GATHER, L, s);
}
void adj_fused(Prec pr, const Launch& L, const Launch& L2, hipStream_t s) {
    if (pr == PREC_F32) launch_fused<float>(L, L2, s);
    else launch_fused<double>(L, L2, s);
}

void winding_numbers(const double* tri, int nf, const double* P, int np, double* partial, int chunks, double* w,
                     hipStream_t s) {
    if (np <= 0) return;
    const int per = chunks > 0 ? (nf + chunks - 1) / chunks : 0;
    const unsigned nbx = (unsigned)((np + 255) / 256);
    hipLaunchKernelGGL(k_winding, dim3(nbx, (unsigned)chunks), dim3(256), 0, s, tri, nf, per, P, np, partial);
    hipLaunchKernelGGL(k_winding_sum, dim3(nbx), dim3(256), 0, s, partial, chunks, np, w);
}

template <typename T>
__global__ void k_permute(const T* src, T* dst, const int* q, int n, int nc, long long ld) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int i = q[j];
    for (int c = 0; c < nc; ++c) dst[c * ld + j] = src[c * ld + i];
}

void permute_planes(Prec pr, const void* src, void* dst, const int* q, int n, int nc, long long ld, hipStream_t s) {
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32)
        hipLaunchKernelGGL(k_permute<float>, dim3(nb), dim3(256), 0, s, (const float*)src, (float*)dst, q, n, nc, ld);
    else
        hipLaunchKernelGGL(k_permute<double>, dim3(nb), dim3(256), 0, s, (const double*)src, (double*)dst, q, n, nc,
                           ld);
}

void permute_ints(const int* src, int* dst, const int* q, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_permute<int>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, q, n, 1, 0LL);
}

__global__ void k_invert_perm(const int* q, int* inv, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) inv[q[j]] = j;
}

void invert_perm(const int* q, int* inv, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_invert_perm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q, inv, n);
}

void fill_zero(Prec pr, void* p, long long n, hipStream_t s) {
    if (n <= 0) return;
    const long long bytes = n * (pr == PREC_F32 ? 4 : 8);
    if (((uintptr_t)p & 15) == 0 && (bytes & 15) == 0) {
        const long long n16 = bytes / 16;
        const long long nb = std::min<long long>((n16 + 255) / 256, 8192);
        hipLaunchKernelGGL(k_fill_zero16, dim3((unsigned)nb), dim3(256), 0, s, (uint4*)p, n16);
        return;
    }
    if (pr == PREC_F32) hipLaunchKernelGGL(k_fill_zero<float>, dim3(ew_blocks(n)), dim3(256), 0, s, (float*)p, n);
    else hipLaunchKernelGGL(k_fill_zero<double>, dim3(ew_blocks(n)), dim3(256), 0, s, (double*)p, n);
}

void import_planes(Prec pr, const void* src, int src_f64, void* dst, const int* perm, int n, int nc, long long ld,
                   hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32) {
        if (src_f64) hipLaunchKernelGGL((k_import<float, double>), dim3(nb), dim3(256), 0, s, (const double*)src, (float*)dst, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_import<float, float>), dim3(nb), dim3(256), 0, s, (const float*)src, (float*)dst, perm, n, nc, ld);
    } else {
        if (src_f64) hipLaunchKernelGGL((k_import<double, double>), dim3(nb), dim3(256), 0, s, (const double*)src, (double*)dst, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_import<double, float>), dim3(nb), dim3(256), 0, s, (const float*)src, (double*)dst, perm, n, nc, ld);
    }
}

void accumulate_carry(Prec pr, const void* src, int src_f64, void* carry, int plane_off, int nc, const int* perm,
                      int n, long long ld, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32) {
        float* d = (float*)carry + plane_off * ld;
        if (src_f64) hipLaunchKernelGGL((k_accumulate<float, double>), dim3(nb), dim3(256), 0, s, (const double*)src, d, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_accumulate<float, float>), dim3(nb), dim3(256), 0, s, (const float*)src, d, perm, n, nc, ld);
    } else {
        double* d = (double*)carry + plane_off * ld;
        if (src_f64) hipLaunchKernelGGL((k_accumulate<double, double>), dim3(nb), dim3(256), 0, s, (const double*)src, d, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_accumulate<double, float>), dim3(nb), dim3(256), 0, s, (const float*)src, d, perm, n, nc, ld);
    }
}

void export_planes(Prec pr, const void* src, void* dst, int dst_f64, const int* perm, int n, int nc, long long ld,
                   hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32) {
        if (dst_f64) hipLaunchKernelGGL((k_export<float, double>), dim3(nb), dim3(256), 0, s, (const float*)src, (double*)dst, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_export<float, float>), dim3(nb), dim3(256), 0, s, (const float*)src, (float*)dst, perm, n, nc, ld);
    } else {
        if (dst_f64) hipLaunchKernelGGL((k_export<double, double>), dim3(nb), dim3(256), 0, s, (const double*)src, (double*)dst, perm, n, nc, ld);
        else hipLaunchKernelGGL((k_export<double, float>), dim3(nb), dim3(256), 0, s, (const double*)src, (float*)dst, perm, n, nc, ld);
    }
}

template <typename T, typename D>
static void launch_export_stress(int model, const void* st, const void* ctl, const void* stat, void* dst,
                                 const int* perm, int n, long long ld, hipStream_t s) {
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (model == 0)
        hipLaunchKernelGGL((k_export_stress<T, D, 0>), dim3(nb), dim3(256), 0, s, (const T*)st, (const T*)ctl,
                           (const T*)stat, (D*)dst, perm, n, ld);
    else
        hipLaunchKernelGGL((k_export_stress<T, D, 1>), dim3(nb), dim3(256), 0, s, (const T*)st, (const T*)ctl,
                           (const T*)stat, (D*)dst, perm, n, ld);
}

void export_stress(Prec pr, int model, const void* st, const void* ctl, const void* stat, void* dst, int dst_f64,
                   const int* perm, int n, long long ld, hipStream_t s) {
    if (n <= 0) return;
    if (pr == PREC_F32) {
        if (dst_f64) launch_export_stress<float, double>(model, st, ctl, stat, dst, perm, n, ld, s);
        else launch_export_stress<float, float>(model, st, ctl, stat, dst, perm, n, ld, s);
    } else {
        if (dst_f64) launch_export_stress<double, double>(model, st, ctl, stat, dst, perm, n, ld, s);
        else launch_export_stress<double, float>(model, st, ctl, stat, dst, perm, n, ld, s);
    }
}

void export_F_total(Prec pr, const void* st, const void* ctl, void* dst, int dst_f64, const int* perm, int n,
                    long long ld, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32) {
        if (dst_f64) hipLaunchKernelGGL((k_export_Ftot<float, double>), dim3(nb), dim3(256), 0, s, (const float*)st, (const float*)ctl, (double*)dst, perm, n, ld);
        else hipLaunchKernelGGL((k_export_Ftot<float, float>), dim3(nb), dim3(256), 0, s, (const float*)st, (const float*)ctl, (float*)dst, perm, n, ld);
    } else {
        if (dst_f64) hipLaunchKernelGGL((k_export_Ftot<double, double>), dim3(nb), dim3(256), 0, s, (const double*)st, (const double*)ctl, (double*)dst, perm, n, ld);
        else hipLaunchKernelGGL((k_export_Ftot<double, float>), dim3(nb), dim3(256), 0, s, (const double*)st, (const double*)ctl, (float*)dst, perm, n, ld);
    }
}

void copy_elems(Prec pr, void* dst, const void* src, long long count, hipStream_t s) {
    if (count <= 0) return;
    if (pr == PREC_F32) hipLaunchKernelGGL(k_copy<float>, dim3(ew_blocks(count)), dim3(256), 0, s, (float*)dst, (const float*)src, count);
    else hipLaunchKernelGGL(k_copy<double>, dim3(ew_blocks(count)), dim3(256), 0, s, (double*)dst, (const double*)src, count);
}

void mass_sums(Prec pr, const void* gm, const void* tm, long long G, double* partial, int nblocks, hipStream_t s) {
    if (pr == PREC_F32) hipLaunchKernelGGL(k_mass_sums<float>, dim3(nblocks), dim3(256), 0, s, (const float*)gm, (const float*)tm, G, partial);
    else hipLaunchKernelGGL(k_mass_sums<double>, dim3(nblocks), dim3(256), 0, s, (const double*)gm, (const double*)tm, G, partial);
}

void mass_loss_nodes(Prec pr, const void* gm, const void* tm, long long G, const double* sums, int, void* dldm,
                     double* partial, int nblocks, hipStream_t s) {
    if (pr == PREC_F32) hipLaunchKernelGGL(k_mass_loss_nodes<float>, dim3(nblocks), dim3(256), 0, s, (const float*)gm, (const float*)tm, G, sums, (float*)dldm, partial);
    else hipLaunchKernelGGL(k_mass_loss_nodes<double>, dim3(nblocks), dim3(256), 0, s, (const double*)gm, (const double*)tm, G, sums, (double*)dldm, partial);
}

void mass_loss_particles(Prec pr, const Launch& L, const void* dldm, const void* tm, hipStream_t s) {
    const unsigned nb = (unsigned)((L.n + 255) / 256);
    if (L.n <= 0) return;
    if (pr == PREC_F32) {
        KP<float> k = make_kp<float>(L, 0);
        if (L.ph.order == 3) hipLaunchKernelGGL((k_mass_loss_particles<float, 3>), dim3(nb), dim3(256), 0, s, k, (const float*)dldm, (const float*)tm);
        else hipLaunchKernelGGL((k_mass_loss_particles<float, 2>), dim3(nb), dim3(256), 0, s, k, (const float*)dldm, (const float*)tm);
    } else {
        KP<double> k = make_kp<double>(L, 0);
        if (L.ph.order == 3) hipLaunchKernelGGL((k_mass_loss_particles<double, 3>), dim3(nb), dim3(256), 0, s, k, (const double*)dldm, (const double*)tm);
        else hipLaunchKernelGGL((k_mass_loss_particles<double, 2>), dim3(nb), dim3(256), 0, s, k, (const double*)dldm, (const double*)tm);
    }
}

void reduce_partials(const double* partial, int nblocks, int stride, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, s, partial, nblocks, stride, out);
}

void volumes(Prec pr, const Launch& L, void* vol, hipStream_t s) {
    if (L.n <= 0) return;
    const unsigned nb = (unsigned)((L.n + 255) / 256);
    if (pr == PREC_F32) {
        KP<float> k = make_kp<float>(L, 0);
        if (L.ph.order == 3) hipLaunchKernelGGL((k_volumes<float, 3>), dim3(nb), dim3(256), 0, s, k, (float*)vol);
        else hipLaunchKernelGGL((k_volumes<float, 2>), dim3(nb), dim3(256), 0, s, k, (float*)vol);
    } else {
        KP<double> k = make_kp<double>(L, 0);
        if (L.ph.order == 3) hipLaunchKernelGGL((k_volumes<double, 3>), dim3(nb), dim3(256), 0, s, k, (double*)vol);
        else hipLaunchKernelGGL((k_volumes<double, 2>), dim3(nb), dim3(256), 0, s, k, (double*)vol);
    }
}

void sqnorm_planes(Prec pr, const void* p, int n, int nc, long long ld, double* partial, int nblocks, hipStream_t s) {
    if (pr == PREC_F32)
        hipLaunchKernelGGL(k_sqnorm<float>, dim3(nblocks), dim3(256), 0, s, (const float*)p, n, nc, ld, partial);
    else
        hipLaunchKernelGGL(k_sqnorm<double>, dim3(nblocks), dim3(256), 0, s, (const double*)p, n, nc, ld, partial);
}

void sqnorm9(Prec pr, const void* p, int n, long long ld, double* partial, int nblocks, hipStream_t s) {
    sqnorm_planes(pr, p, n, 9, ld, partial, nblocks, s);
}

void adam(Prec pr, const void* dLdF, void* dFc, void* mom, void* vec, void* vmax, int n, long long ld,
          const double* sqn, double alpha, double beta1, double beta2, double eps, int timestep, hipStream_t s) {
    if (n <= 0) return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (pr == PREC_F32)
        hipLaunchKernelGGL(k_adam<float>, dim3(nb), dim3(256), 0, s, (const float*)dLdF, (float*)dFc, (float*)mom, (float*)vec, (float*)vmax, n, ld, sqn, (float)alpha, (float)beta1, (float)beta2, (float)eps, timestep);
    else
        hipLaunchKernelGGL(k_adam<double>, dim3(nb), dim3(256), 0, s, (const double*)dLdF, (double*)dFc, (double*)mom, (double*)vec, (double*)vmax, n, ld, sqn, alpha, beta1, beta2, eps, timestep);
}

// ---- particle ordering on the device ----------------------------------------
template <typename S, bool SOA>
DMPM_DI double order_pos(const S* x, long long ld, int i, int a) {
    return SOA ? (double)x[(long long)a * ld + i] : (double)x[3 * (long long)i + a];
}

template <typename S, bool SOA>
__global__ void k_order_keys(const S* x, long long ld, int n, double mn0, double mn1, double mn2, double dx,
                             double shift, unsigned* keys, int* vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mn[3] = {mn0, mn1, mn2};
    uint32_t c[3];
    for (int a = 0; a < 3; ++a) {
        const double f = (order_pos<S, SOA>(x, ld, i, a) - mn[a]) / dx;
        long long ci = (long long)floor(f - shift) + 2;
        ci = ci < 0 ? 0 : (ci > 1023 ? 1023 : ci);
        c[a] = (uint32_t)ci;
    }
    keys[i] = hilbert3(c[0], c[1], c[2]);
    vals[i] = i;
}

// Base cell of the j-th particle in sorted order.
template <typename S, bool SOA>
__global__ void k_order_cells(const S* x, long long ld, int n, const int* perm, double mn0, double mn1, double mn2,
                              double dx, double shift, int* cell) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double mn[3] = {mn0, mn1, mn2};
    const int p = perm[j];
    for (int a = 0; a < 3; ++a) {
        const double f = (order_pos<S, SOA>(x, ld, p, a) - mn[a]) / dx;
        cell[3 * (long long)j + a] = (int)fmax(-1e9, fmin(1e9, floor(f - shift)));
    }
}

// Block cuts inside fixed chunks of B (at most 256) particles, one wave per
// chunk.  From a block's first particle s the wave forms, for every later
// particle of the chunk, the box of the base cells of s..that particle in
// each of the SETS base-cell sets (cell + q * 3n holds set q: the sort
// layer's cells, then other layers') -- lane l holds particles s + 4l ..
// s + 4l + 3, its own running box and the boxes of the lanes below through
// a wave scan -- and the block ends before the first particle whose
// stencil box exceeds `split` nodes in one of the sets (boxes only grow, so
// every later particle fails too); the next block starts there.  Every block
// therefore holds at most B particles and fits in every set (a lone
// particle always makes a block).  The chunk's block starts go to
// cut[cB ...], their count to cnt[c].
template <int SETS>
__global__ void __launch_bounds__(64) k_chunk_cuts(const int* cell, int n, int B, int KS, long long split, int* cut,
                                                   int* cnt) {
    const int lane = threadIdx.x;
    const long long i0 = (long long)blockIdx.x * B;
    const int m = (int)min((long long)B, (long long)n - i0);
    int k = 0;
    for (int s = 0; s < m;) {  // s, k: the same in every lane
        // running boxes of this lane's particles, and their cells
        int lo[4][SETS][3], hi[4][SETS][3];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int p = s + 4 * lane + i;
#pragma unroll
            for (int q = 0; q < SETS; ++q)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const int v = p < m ? cell[(long long)q * 3 * n + 3 * (i0 + p) + a] : 0;
                    const int l0 = i == 0 ? INT_MAX : lo[i - 1][q][a], h0 = i == 0 ? INT_MIN : hi[i - 1][q][a];
                    lo[i][q][a] = p < m ? min(l0, v) : l0;
                    hi[i][q][a] = p < m ? max(h0, v) : h0;
                }
        }
        // boxes of the lanes below (exclusive wave scan of the lane boxes)
        int elo[SETS][3], ehi[SETS][3];
#pragma unroll
        for (int q = 0; q < SETS; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int vl = lo[3][q][a], vh = hi[3][q][a];
#pragma unroll
                for (int d = 1; d < 64; d *= 2) {
                    const int ul = __shfl_up(vl, d), uh = __shfl_up(vh, d);
                    if (lane >= d) {
                        vl = min(vl, ul);
                        vh = max(vh, uh);
                    }
                }
                const int xl = __shfl_up(vl, 1), xh = __shfl_up(vh, 1);
                elo[q][a] = lane == 0 ? INT_MAX : xl;
                ehi[q][a] = lane == 0 ? INT_MIN : xh;
            }
        // this lane's first particle whose box from s does not fit (4: none)
        int first = 4;
#pragma unroll
        for (int i = 3; i >= 0; --i) {
            const int p = s + 4 * lane + i;
            bool fits = true;
#pragma unroll
            for (int q = 0; q < SETS; ++q) {
                long long vol = 1;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    vol *= (long long)max(ehi[q][a], hi[i][q][a]) - min(elo[q][a], lo[i][q][a]) + KS;
                fits = fits && vol <= split;
            }
            if (p < m && p > s && !fits) first = i;
        }
        const unsigned long long bad = __ballot(first < 4);
        const int end = bad ? s + 4 * (__ffsll((long long)bad) - 1) + __shfl(first, __ffsll((long long)bad) - 1) : m;
        if (lane == 0) cut[i0 + k] = (int)i0 + s;
        ++k;
        s = end;
    }
    if (lane == 0) cnt[blockIdx.x] = k;
}

// Block starts of all chunks in order (off: exclusive prefix sum of cnt),
// the end sentinel n and the block count.
__global__ void k_chunk_starts(const int* cut, const int* cnt, const int* off, int nchunk, int n, int B, int* starts,
                               int* nblk) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunk) return;
    const int o = off[c], k = cnt[c];
    for (int q = 0; q < k; ++q) starts[o + q] = cut[(long long)c * B + q];
    if (c == nchunk - 1) {
        starts[o + k] = n;
        *nblk = o + k;
    }
}

// One wave per block: its runs of equal base cells placed row by row (16
// lanes), the longest runs first, each into the current row's remaining room
// when one fits (the longest that fits, the earliest of equal length), else
// the longest left.  The lanes find the runs (ballots of run heads) and rank
// them by length; one lane places them, taking runs from per-length queues;
// the lanes move the particles.  Blocks hold at most 256 particles.
__global__ void __launch_bounds__(64) k_pack_rows(const int* cell, const int* starts, const int* perm_in,
                                                  int* perm_out) {
    __shared__ unsigned short s_at[257];     // run starts (block-relative); s_at[R] = block length
    __shared__ unsigned short s_sorted[256]; // runs, longest first (stable)
    __shared__ unsigned short s_dest[256];   // block-relative destination of each run
    __shared__ unsigned short s_next[257];   // per run length: next unplaced entry of s_sorted
    __shared__ unsigned short s_end[257];    // per run length: end of its entries in s_sorted
    const int s0 = starts[blockIdx.x], len = starts[blockIdx.x + 1] - s0;
    const int lane = threadIdx.x;
    for (int L = lane; L < 257; L += 64) s_next[L] = s_end[L] = 0;
    int run_of[4];
    int R = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int r = k * 64 + lane;
        bool head = false;
        if (r < len) {
            const long long j = s0 + r;
            head = r == 0 || cell[3 * j] != cell[3 * j - 3] || cell[3 * j + 1] != cell[3 * j - 2] ||
                   cell[3 * j + 2] != cell[3 * j - 1];
        }
        const unsigned long long m = __ballot(head);
        const int before = __popcll(m & ((1ull << lane) - 1));
        if (head) s_at[R + before] = (unsigned short)r;
        run_of[k] = R + before + (head ? 0 : -1);
        R += __popcll(m);
    }
    if (lane == 0) s_at[R] = (unsigned short)len;
    __syncthreads();
    for (int r = lane; r < R; r += 64) {
        const int L = s_at[r + 1] - s_at[r];
        int rank = 0;
        for (int q = 0; q < R; ++q) {
            const int Lq = s_at[q + 1] - s_at[q];
            rank += Lq > L || (Lq == L && q < r);
        }
        s_sorted[rank] = (unsigned short)r;
    }
    __syncthreads();
    if (lane == 0 && R > 0) {
        for (int i = 0; i < R; ++i) {
            const int L = s_at[s_sorted[i] + 1] - s_at[s_sorted[i]];
            if (s_end[L] == 0) s_next[L] = (unsigned short)i;
            s_end[L] = (unsigned short)(i + 1);
        }
        int maxL = s_at[s_sorted[0] + 1] - s_at[s_sorted[0]];
        int out = 0;
        for (int placed = 0; placed < R; ++placed) {
            const int room = 16 - out % 16;
            int pick = 0;
            for (int L = min(room, maxL); L >= 1; --L)
                if (s_next[L] < s_end[L]) {
                    pick = L;
                    break;
                }
            if (pick == 0) {
                while (s_next[maxL] >= s_end[maxL]) --maxL;
                pick = maxL;
            }
            const int r = s_sorted[s_next[pick]++];
            s_dest[r] = (unsigned short)out;
            out += pick;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int r = k * 64 + lane;
        if (r < len) perm_out[s0 + s_dest[run_of[k]] + (r - s_at[run_of[k]])] = perm_in[s0 + r];
    }
}

namespace {
struct OrderLayout {
    unsigned *kin, *kout;
    int *vin, *vout, *cell, *cut, *cnt, *off, *nblk;
    void* temp;
    size_t temp_bytes;
};
// temporary storage of the radix sort and of the chunk scan (used one after
// the other)
size_t order_cub_bytes(int n) {
    size_t sort = 0, scan = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort, (const unsigned*)nullptr, (unsigned*)nullptr,
                                             (const int*)nullptr, (int*)nullptr, n, 0, 30, (hipStream_t)0);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int*)nullptr, (int*)nullptr, std::max(n, 1),
                                           (hipStream_t)0);
    return std::max(sort, scan);
}
OrderLayout order_layout(void* base, int n, int sets) {
    OrderLayout L{};
    char* p = (char*)base;
    auto take = [&](size_t bytes) {
        char* q = p;
        p += (bytes + 255) & ~size_t(255);
        return q;
    };
    const size_t nn = (size_t)std::max(n, 1);
    L.kin = (unsigned*)take(4 * nn);
    L.kout = (unsigned*)take(4 * nn);
    L.vin = (int*)take(4 * nn);
    L.vout = (int*)take(4 * nn);
    L.cell = (int*)take(12 * nn * (size_t)sets);
    L.cut = (int*)take(4 * nn);
    L.cnt = (int*)take(4 * nn);
    L.off = (int*)take(4 * nn);
    L.nblk = (int*)take(4);
    L.temp_bytes = order_cub_bytes(n);
    L.temp = take(L.temp_bytes);
    return L;
}
}  // namespace

size_t order_scratch_bytes(int n, int ncut) {
    return (44 + 12 * (size_t)ncut) * (size_t)std::max(n, 1) + order_cub_bytes(n) + 11 * 256;
}

int order_particles_device(int order, int block, int split_nodes, const double grid_min[3], double dx,
                           const void* x, int x_f64, int soa, long long ld, int n, int* perm, int* starts,
                           void* scratch, size_t scratch_bytes, hipStream_t s, const void* const* xcut, int ncut) {
    if (n <= 0) return 0;
    if (ncut < 0 || ncut > kOrderCutSets - 1) throw std::runtime_error("particle ordering: too many cut sets");
    if (scratch_bytes < order_scratch_bytes(n, ncut)) throw std::runtime_error("particle ordering: scratch too small");
    OrderLayout L = order_layout(scratch, n, 1 + ncut);
    const double shift = order == 2 ? 0.5 : 0.0;
    const unsigned nb = (unsigned)((n + 255) / 256);
#define ORDER_LAUNCH(KERNEL, X, ...)                                                                        \
    do {                                                                                                    \
        if (x_f64) {                                                                                        \
            if (soa) hipLaunchKernelGGL((KERNEL<double, true>), dim3(nb), dim3(256), 0, s, (const double*)(X), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<double, false>), dim3(nb), dim3(256), 0, s, (const double*)(X), __VA_ARGS__);   \
        } else {                                                                                            \
            if (soa) hipLaunchKernelGGL((KERNEL<float, true>), dim3(nb), dim3(256), 0, s, (const float*)(X), __VA_ARGS__);   \
            else hipLaunchKernelGGL((KERNEL<float, false>), dim3(nb), dim3(256), 0, s, (const float*)(X), __VA_ARGS__);     \
        }                                                                                                   \
    } while (0)
    ORDER_LAUNCH(k_order_keys, x, ld, n, grid_min[0], grid_min[1], grid_min[2], dx, shift, L.kin, L.vin);
    size_t tb = L.temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(L.temp, tb, L.kin, L.kout, L.vin, L.vout, n, 0, 30, s) != hipSuccess)
        throw std::runtime_error("particle ordering: radix sort failed");
    const long long split = split_nodes == 0 ? kDefaultSplitNodes : split_nodes;
    if (split <= 0) {
        if (hipMemcpyAsync(perm, L.vout, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, s) != hipSuccess)
            throw std::runtime_error("particle ordering: copy failed");
        return 0;
    }
    // base cells in the sorted order: of x (the runs and the cuts), then of
    // each further set (the cuts only)
    ORDER_LAUNCH(k_order_cells, x, ld, n, L.vout, grid_min[0], grid_min[1], grid_min[2], dx, shift, L.cell);
    for (int q = 0; q < ncut; ++q)
        ORDER_LAUNCH(k_order_cells, xcut[q], ld, n, L.vout, grid_min[0], grid_min[1], grid_min[2], dx, shift,
                     L.cell + (size_t)(q + 1) * 3 * n);
#undef ORDER_LAUNCH
    const int B = block > 0 ? block : 256;
    if (B > 256) throw std::runtime_error("particle ordering: blocks hold at most 256 particles");
    const int nchunk = (int)(((long long)n + B - 1) / B);
    const unsigned nbc = (unsigned)((nchunk + 255) / 256);
    {
        const int KS = order == 2 ? 3 : 4;
        static_assert(kOrderCutSets == 3, "one k_chunk_cuts instance per set count");
        if (ncut == 0)
            hipLaunchKernelGGL(k_chunk_cuts<1>, dim3((unsigned)nchunk), dim3(64), 0, s, L.cell, n, B, KS, split, L.cut,
                               L.cnt);
        else if (ncut == 1)
            hipLaunchKernelGGL(k_chunk_cuts<2>, dim3((unsigned)nchunk), dim3(64), 0, s, L.cell, n, B, KS, split, L.cut,
                               L.cnt);
        else
            hipLaunchKernelGGL(k_chunk_cuts<3>, dim3((unsigned)nchunk), dim3(64), 0, s, L.cell, n, B, KS, split, L.cut,
                               L.cnt);
    }
    size_t tsb = L.temp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(L.temp, tsb, L.cnt, L.off, nchunk, s) != hipSuccess)
        throw std::runtime_error("particle ordering: block scan failed");
    hipLaunchKernelGGL(k_chunk_starts, dim3(nbc), dim3(256), 0, s, L.cut, L.cnt, L.off, nchunk, n, B, starts,
                       L.nblk);
    int nblk = 0;
    if (hipMemcpyAsync(&nblk, L.nblk, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        throw std::runtime_error("particle ordering: block count read failed");
    if (nblk > 0)
        hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)nblk), dim3(64), 0, s, L.cell, starts, L.vout, perm);
    return nblk;
}

// ---- deterministic mode ----------------------------------------------------
size_t det_scratch_bytes(int n, const int* dims, int esize) {
    const long long ncells = (long long)(dims[0] + 8) * (dims[1] + 8) * (dims[2] + 8);
    size_t temp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 32, (hipStream_t)0);
    const long long pld = (n + 63) / 64 * 64;
    // 8 chunks, each rounded up to 256 bytes by det_layout
    return 4 * (size_t)pld * 4 + 2 * (size_t)ncells * 4 + 13 * (size_t)pld * esize + temp + 8 * 256;
}

struct DetLayout {
    unsigned *kin, *kout;
    int *iin, *iout, *cstart, *cend;
    void* pay;
    void* temp;
    size_t temp_bytes;
    long long pld, ncells;
};

static DetLayout det_layout(void* scratch, int n, const int* dims, int esize) {
    DetLayout d;
    d.ncells = (long long)(dims[0] + 8) * (dims[1] + 8) * (dims[2] + 8);
    d.pld = (n + 63) / 64 * 64;
    char* p = (char*)scratch;
    auto take = [&](size_t bytes) {
        void* r = p;
        p += (bytes + 255) / 256 * 256;
        return r;
    };
    d.kin = (unsigned*)take(d.pld * 4);
    d.kout = (unsigned*)take(d.pld * 4);
    d.iin = (int*)take(d.pld * 4);
    d.iout = (int*)take(d.pld * 4);
    d.cstart = (int*)take(d.ncells * 4);
    d.cend = (int*)take(d.ncells * 4);
    d.pay = take(13 * d.pld * esize);
    d.temp_bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, d.temp_bytes, d.kin, d.kout, d.iin, d.iout, n, 0, 32, (hipStream_t)0);
    d.temp = take(d.temp_bytes);
    return d;
}

template <typename T>
static void det_sort_ranges(const Launch& L, DetLayout& d, hipStream_t s) {
    // empty cells: start = end = 0 (32-bit zero fill)
    fill_zero(PREC_F32, d.cstart, d.ncells, s);
    fill_zero(PREC_F32, d.cend, d.ncells, s);
    (void)hipcub::DeviceRadixSort::SortPairs(d.temp, d.temp_bytes, d.kin, d.kout, d.iin, d.iout, L.n, 0, 32, s);
    hipLaunchKernelGGL(k_det_cell_ranges, dim3((L.n + 255) / 256), dim3(256), 0, s, d.kout, L.n, d.cstart, d.cend);
}

template <typename T>
static void launch_det_p2g(const Launch& L, bool mass_only, void* scratch, hipStream_t s) {
    if (L.n <= 0) return;
    KP<T> k = make_kp<T>(L, 0);
    DetLayout d = det_layout(scratch, L.n, L.ph.dims, (int)sizeof(T));
    const unsigned nbp = (unsigned)((L.n + 255) / 256);
    const long long G = (long long)L.ph.dims[0] * L.ph.dims[1] * L.ph.dims[2];
    const unsigned nbg = (unsigned)((G + 255) / 256);
    T* pay = (T*)d.pay;
    const int o = L.ph.order, md = L.ph.model;
#define DP_CASE(O, M, MO) hipLaunchKernelGGL((k_det_prep_p2g<T, O, M, MO>), dim3(nbp), dim3(256), 0, s, k, d.kin, d.iin, pay, d.pld)
    if (mass_only) { if (o == 3) DP_CASE(3, 0, true); else DP_CASE(2, 0, true); }
    else if (o == 3) { if (md == 0) DP_CASE(3, 0, false); else DP_CASE(3, 1, false); }
    else { if (md == 0) DP_CASE(2, 0, false); else DP_CASE(2, 1, false); }
#undef DP_CASE
    det_sort_ranges<T>(L, d, s);
#define DG_CASE(O, MO) hipLaunchKernelGGL((k_det_gather_p2g<T, O, MO>), dim3(nbg), dim3(256), 0, s, k, d.cstart, d.cend, d.iout, (const T*)pay, d.pld)
    if (mass_only) { if (o == 3) DG_CASE(3, true); else DG_CASE(2, true); }
    else { if (o == 3) DG_CASE(3, false); else DG_CASE(2, false); }
#undef DG_CASE
}

template <typename T>
static void launch_det_adj(const Launch& L, void* scratch, hipStream_t s) {
    if (L.n <= 0) return;
    KP<T> k = make_kp<T>(L, 0);
    DetLayout d = det_layout(scratch, L.n, L.ph.dims, (int)sizeof(T));
    const unsigned nbp = (unsigned)((L.n + 255) / 256);
    const long long G = (long long)L.ph.dims[0] * L.ph.dims[1] * L.ph.dims[2];
    const unsigned nbg = (unsigned)((G + 255) / 256);
    T* pay = (T*)d.pay;
    const bool ref = L.ph.adjoint == 1;
    const int o = L.ph.order;
#define DA_CASE(O, R) hipLaunchKernelGGL((k_det_prep_adj<T, O, R>), dim3(nbp), dim3(256), 0, s, k, d.kin, d.iin, pay, d.pld)
    if (o == 3) { if (ref) DA_CASE(3, true); else DA_CASE(3, false); }
    else { if (ref) DA_CASE(2, true); else DA_CASE(2, false); }
#undef DA_CASE
    det_sort_ranges<T>(L, d, s);
    if (o == 3) hipLaunchKernelGGL((k_det_gather_adj<T, 3>), dim3(nbg), dim3(256), 0, s, k, d.cstart, d.cend, d.iout, (const T*)pay, d.pld);
    else hipLaunchKernelGGL((k_det_gather_adj<T, 2>), dim3(nbg), dim3(256), 0, s, k, d.cstart, d.cend, d.iout, (const T*)pay, d.pld);
}

void det_p2g(Prec pr, const Launch& L, bool mass_only, void* scratch, hipStream_t s) {
    if (pr == PREC_F32) launch_det_p2g<float>(L, mass_only, scratch, s);
    else launch_det_p2g<double>(L, mass_only, scratch, s);
}

void det_adj_scatter(Prec pr, const Launch& L, void* scratch, hipStream_t s) {
    if (pr == PREC_F32) launch_det_adj<float>(L, scratch, s);
    else launch_det_adj<double>(L, scratch, s);
}

#endif  // DMPM_IN(0)

}  // namespace dmpm
