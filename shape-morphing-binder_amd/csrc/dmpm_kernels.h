// Host-side launchers for the MPM kernels (dmpm_kernels.hip).
//
// Device data layout (struct-of-arrays, "planes"):
//   layer state  : 24 planes of `ld` elements: x0..2, v0..2, F00..F22, C00..C22
//   statics      : 4 planes: m, vol, lam, mu
//   control dFc  : 9 planes
//   adjoint carry: 24 planes, same order as the state (dL/dx, dL/dv, dL/dF, dL/dC)
//   grid         : node (i,j,k) at g = i*ny*nz + j*nz + k (the reference's
//                  order, Grid.h:89-94), its values interleaved at [4g, 4g+4):
//                  m, p0..2 (forward), dL/dv_i 0..2 and an unused slot
//                  (adjoint); `gld` >= nx*ny*nz nodes.  A mass-only P2G (the
//                  end-layer loss, p2g, point-cloud volumes) writes a plane of
//                  one mass per node.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <cstddef>
#include <cstdint>

namespace dmpm {

enum Prec { PREC_F32 = 0, PREC_F64 = 1 };

// Plane offsets inside a layer state / adjoint carry.
enum { PX = 0, PV = 3, PF = 6, PC = 15, NSTATE = 24 };
enum { SM = 0, SVOL = 1, SLAM = 2, SMU = 3, NSTAT = 4 };

struct Phys {
    int dims[3];
    double dx, grid_min[3];
    double dt, drag, smoothing, zeta, jmin, f_ext[3];
    int order, model, trust_region, wall_cells, adjoint;
};

// Everything one kernel launch may need; pointers are device pointers of the
// precision in use (float or double), passed as void*.
struct Launch {
    Phys ph;
    int n;          // particles
    long long ld;   // particle plane stride (elements)
    long long gld;  // grid nodes allocated (>= nx*ny*nz, a multiple of 64)
    const void* st;    // layer state (24 planes)
    const void* stat;  // statics (4 planes)
    const void* ctl;   // dFc (9 planes) or null
    void* out;         // next layer state (24 planes)
    void* grid;        // m, p0, p1, p2 per node (4 values each); mass-only P2G: m per node
    void* ggv;         // dL/dv_i per node (4 values each, the last unused)
    const void* adj1;  // adjoint carry of layer t+1
    void* adj0;        // adjoint carry of layer t (output)
    void* gctl;        // dL/d dFc of layer t (output, 9 planes) or null
    const void* ctl1;  // fused step: dFc of layer t+1 or null
    void* grid1;       // fused step: grid t+1 (as grid, zeroed by the caller)
    // adjoint: state of layer t+1 when it is known to be the forward step of
    // `st` under the current control and parameters (its v and C are then
    // read instead of recomputed by a G2P from `grid`), else null
    const void* nxt;
    // rows in another storage order: particle i's new state (forward step)
    // or adjoint carry (unfused adjoint gather) is written to row opos[i]
    // instead of i (a step between segments of two orders), or null
    const int* opos;
    // particle blocks of the window kernels: block b holds particles
    // [bst[b], bst[b+1]) (nblk blocks of at most `block` particles, cut where
    // a block's stencil box would grow too large); null: consecutive blocks of
    // `block` particles
    const int* bst;
    int nblk;
    unsigned long long* wstat;  // window counters: [0] blocks in LDS, [1] blocks falling back (or null)
    // fused step and fused adjoint: a buffer that a later launch on the
    // stream accumulates into, cleared by this launch's blocks after their
    // own work (16-byte aligned, zero_bytes a multiple of 16), or null
    void* zero;
    long long zero_bytes;
    int lds_bytes;     // LDS budget per block for the window
    int block;         // threads per block (multiple of 64)
};

// Generalised winding numbers w[np] of points P (np x 3) w.r.t. the
// triangles tri (nf x 9: the three vertices of each), all device memory;
// the triangles are split into `chunks` ranges summed through partial
// (chunks x np doubles).
void winding_numbers(const double* tri, int nf, const double* P, int np, double* partial, int chunks, double* w,
                     hipStream_t s);

// ---- batched launches: one kernel launch over many instances ---------------
// Operations of a recorded episode.
enum OpKind {
    OP_ZERO = 0,     // dst[0..n) = 0 (n elements)
    OP_COPY,         // dst[0..n) = src[0..n)
    OP_ACCUM,        // dst[0..n) += src[0..n)
    OP_P2G,          // Launch: P2G of m and p
    OP_P2G_MASS,     // Launch: P2G of m only
    OP_G2P,
    OP_STEP,
    OP_ADJ_SCATTER,
    OP_ADJ_GATHER,
    OP_ADJ_FUSED,    // Launch of step t (gather) + Launch of step t-1 (scatter)
};
inline bool op_is_kernel(int kind) { return kind >= OP_P2G; }
// Parameter blocks one instance of kernel op `kind` takes (2 for the fused
// adjoint, 1 otherwise).
inline int kp_count(int kind) { return kind == OP_ADJ_FUSED ? 2 : 1; }
// Size of one kernel parameter block for precision pr, and the block(s) of
// kernel op `kind` for launch L (and L2, the second launch of a fused op),
// written to host memory dst.
size_t kp_bytes(Prec pr);
void build_kp(Prec pr, int kind, const Launch& L, const Launch* L2, void* dst);
// Thread blocks one instance takes (its block table's count, or n / block
// rounded up without one).
int instance_blocks(const Launch& L);
// One launch of kernel op `kind` for `count` instances: tab (device) holds
// their parameter blocks, bstart (device, count + 1 entries) the prefix sums
// of their block counts, total = bstart[count].  The instances share the
// precision, physics templates (order, model, adjoint mode), block size and
// LDS budget.
void launch_batch(Prec pr, int kind, const Phys& ph, int lds_bytes, int block, const void* tab, const int* bstart,
                  int count, int total, hipStream_t s);
// Element ranges of an OP_ZERO / OP_COPY / OP_ACCUM over many buffers (one
// launch): segs (device) holds `count` ranges; n counts elements of pr and
// max_n is the largest n.  Zero and copy ranges must be 16-byte aligned and a
// multiple of 16 bytes long.
struct Seg {
    void* dst;
    const void* src;
    long long n;
};
void launch_segs(Prec pr, int kind, const Seg* segs, int count, long long max_n, hipStream_t s);

// Position of a cell along a 3-D Hilbert curve over a 1024^3 lattice (10
// bits per axis), after J. Skilling, "Programming the Hilbert curve" (AIP
// Conf. Proc. 707, 2004): the coordinates are transformed in place into the
// curve's "transposed" index, whose bits are then interleaved.  Unlike a
// Morton (Z-order) curve the Hilbert curve has no jumps -- consecutive cells
// are face neighbours -- so any run of ~60 cells lies inside two adjacent
// 4x4x4 sub-cubes, which bounds the block windows of the particle kernels.
__host__ __device__ inline uint32_t hilbert3(uint32_t x0, uint32_t x1, uint32_t x2) {
    const int bits = 10;
    uint32_t X[3] = {x0 & 1023u, x1 & 1023u, x2 & 1023u};
    const uint32_t M = 1u << (bits - 1);
    for (uint32_t Q = M; Q > 1; Q >>= 1) {
        const uint32_t P = Q - 1;
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;
            } else {
                const uint32_t t = (X[0] ^ X[i]) & P;
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
    uint32_t t = 0;
    for (uint32_t Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    for (int i = 0; i < 3; ++i) X[i] ^= t;
    uint32_t h = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
    return h;
}

// Storage order and thread blocks of n particles, computed on the device
// from positions that are either the caller's (n,3) array (soa = 0) or three
// planes of stride ld (soa = 1), float or double (x_f64):
//  * the particles sorted (stable radix sort) by the Hilbert index of their
//    stencil base cell, floor(x/dx - 1/2) for the quadratic B-spline and
//    floor(x/dx) - 1 for the cubic one (clamped to the 1024^3 lattice), so
//    that particles sharing a stencil are neighbours and the scatters can
//    merge their sums across lanes;
//  * thread blocks of up to `block` (at most 256) consecutive particles:
//    every chunk of `block` consecutive particles is cut before each
//    particle that would grow the box of the current block's stencils
//    beyond split_nodes grid nodes (0: kDefaultSplitNodes; negative: no cuts, consecutive
//    blocks), at x and at each of the `ncut` (at most kOrderCutSets - 1)
//    further position sets xcut[] of the same layout -- other layers of an
//    episode, so that the blocks' windows stay within the bound as the body
//    deforms;
//  * inside each block, its runs of equal base cells re-ordered so that few
//    cross a 16-lane row (best fit: each row takes the longest remaining run
//    that fits its free lanes, else the longest remaining run).
// perm (n ints) receives the source index of each stored particle and
// starts (n + 1 ints of room) the blocks' first particles and n; returns the
// block count (0 without cuts).  Synchronises the stream (the count).
constexpr int kOrderCutSets = 3;
// Default block cut: the largest stencil box whose window, padded to odd
// strides, fits every kernel's share of the default LDS budget (the fused
// adjoint's gather window, 28 B per node in 24 KiB with FP32 state, is the
// tightest).  Fewer cuts mean fewer part-filled blocks.  On the 1M-particle,
// 300-step workload, 704 against 512 nodes: 61.0 against 61.4 ms per
// episode in 13 segment orders, 65.5 against 68.0 ms in one order, and under
// the reference physics (cubic stencil, whose boxes are larger) 120.5
// against 155.4 ms; no window falls back to global atomics (at 768 nodes a
// few do).
constexpr int kDefaultSplitNodes = 704;
size_t order_scratch_bytes(int n, int ncut = 0);
int order_particles_device(int order, int block, int split_nodes, const double grid_min[3], double dx,
                           const void* x, int x_f64, int soa, long long ld, int n, int* perm, int* starts,
                           void* scratch, size_t scratch_bytes, hipStream_t s, const void* const* xcut = nullptr,
                           int ncut = 0);

// Record the message dmpm_last_error() returns (for library parts outside
// the runtime's own error guard).
void set_last_error(const std::string& msg);

// Forward: P2G of layer `st` (stress computed inside) into `grid`.
// mass_only: scatter m only (the dt = 0 P2G used by the loss and volumes).
void p2g(Prec pr, const Launch& L, bool mass_only, hipStream_t s);
// Forward: G2P + particle update, `st` + grid -> `out`.
void g2p(Prec pr, const Launch& L, hipStream_t s);
// Fused forward step: G2P of `st` from `grid`, update -> `out`, then the
// stress and P2G of `out` (with control ctl1) into `grid1`.
void step(Prec pr, const Launch& L, hipStream_t s);
// Adjoint phases 1+2: scatter dL/dv_i into ggv.
void adj_scatter(Prec pr, const Launch& L, hipStream_t s);
// Adjoint phases 3+4 (+ the stress VJP): writes adj0 and gctl.
void adj_gather(Prec pr, const Launch& L, hipStream_t s);
// adj_gather of step t (L) and the adj_scatter of step t-1 (L2, into L2.ggv,
// a different grid from L.ggv) in one launch; the carry of layer t stays in
// registers between the two and C' of step t-1 is C of layer t.  Only for
// steps whose adjoint reads C' from the next layer (no trust region, or the
// reference adjoint) and when layer t is the forward step of layer t-1.
void adj_fused(Prec pr, const Launch& L, const Launch& L2, hipStream_t s);

// Deterministic two-pass mode: scratch size for n particles on a grid of
// `dims` nodes (element size esize), P2G and adjoint scatter that write every
// node (no clear needed) in an order fixed by a stable sort of particles by
// base cell.
size_t det_scratch_bytes(int n, const int* dims, int esize);
void det_p2g(Prec pr, const Launch& L, bool mass_only, void* scratch, hipStream_t s);
void det_adj_scatter(Prec pr, const Launch& L, void* scratch, hipStream_t s);

// Fill `n` elements of `p` with zero.
void fill_zero(Prec pr, void* p, long long n, hipStream_t s);

// AoS user arrays <-> internal planes with permutation:
//   import: dst[c*ld + i] = src[perm[i]*nc + c]   (src float or double)
//   export: dst[perm[i]*nc + c] = src[c*ld + i]
void import_planes(Prec pr, const void* src, int src_f64, void* dst, const int* perm, int n, int nc,
                   long long ld, hipStream_t s);
void export_planes(Prec pr, const void* src, void* dst, int dst_f64, const int* perm, int n, int nc,
                   long long ld, hipStream_t s);
// export of F + dFc
// PK1 stress P(F + dFc) of every particle (model 0: fixed-corotated, 1:
// Neo-Hookean), 9 values per particle in caller order.
void export_stress(Prec pr, int model, const void* st, const void* ctl, const void* stat, void* dst, int dst_f64,
                   const int* perm, int n, long long ld, hipStream_t s);
void export_F_total(Prec pr, const void* st, const void* ctl, void* dst, int dst_f64, const int* perm, int n,
                    long long ld, hipStream_t s);
// Reorder particle planes: dst[c*ld + j] = src[c*ld + q[j]] (c < nc, j < n).
void permute_planes(Prec pr, const void* src, void* dst, const int* q, int n, int nc, long long ld, hipStream_t s);
// dst[j] = src[q[j]] for n ints.
void permute_ints(const int* src, int* dst, const int* q, int n, hipStream_t s);
// inv[q[j]] = j for a permutation q of 0..n-1
void invert_perm(const int* q, int* inv, int n, hipStream_t s);
// Copy between same-layout plane buffers (`count` elements).
void copy_elems(Prec pr, void* dst, const void* src, long long count, hipStream_t s);

// EndLayerMassLoss node pass: per-block partial sums of (c, t) -> partial[2*b]
void mass_sums(Prec pr, const void* grid_m, const void* target_m, long long G, double* partial, int nblocks,
               hipStream_t s);
// node pass 2: dLdm (written into `dldm`) and per-block loss partials
void mass_loss_nodes(Prec pr, const void* grid_m, const void* target_m, long long G, const double* sums,
                     int nsum_blocks, void* dldm, double* partial, int nblocks, hipStream_t s);
// particle pass: writes the last layer's adjoint carry (dL/dx only, rest 0)
void mass_loss_particles(Prec pr, const Launch& L, const void* dldm, const void* target_m, hipStream_t s);
// deterministic final reduction of `nblocks` partials (stride `stride`) into out[0..stride)
void reduce_partials(const double* partial, int nblocks, int stride, double* out, hipStream_t s);

// CalculatePointCloudVolumes particle pass: vol_p = m_p dx^3 / sum_i w m_i
void volumes(Prec pr, const Launch& L, void* vol_out, hipStream_t s);
}  // namespace dmpm

#include "../../include/dmpm.h"

namespace dmpm {
// P2G of the masses of n particles (device arrays x (n,3) and m (n) of
// precision pr, in the caller's order; no block table, so each thread block
// takes 256 consecutive particles) added into the grid's m plane grid_m
// (ForwardSimulation.cpp:234-237 adds onto the Grid it is given), then, when
// vol is non-null, vol_p = m_p dx^3 / sum_i w_ip m_i (:238-270).  `planes`
// is device scratch of 4 planes of ld >= n elements (x0..x2, m), fully
// written before it is read.  Queued on st; nothing is allocated.
void mass_volume_pass(Prec pr, const dmpm_config& cfg, int n, long long ld, const void* x, const void* m,
                      void* planes, void* grid_m, void* vol, hipStream_t st);

// Squared Frobenius norm partials of 9 planes (dL/dF): partial[b]
void sqnorm9(Prec pr, const void* planes, int n, long long ld, double* partial, int nblocks, hipStream_t s);
// Sum of squares partials of nc planes of n elements (stride ld) of
// precision pr: partial[b].  A flat array of n values is one plane.
void sqnorm_planes(Prec pr, const void* planes, int n, int nc, long long ld, double* partial, int nblocks,
                   hipStream_t s);
// Descend_Adam element update on 9-plane arrays.
void adam(Prec pr, const void* dLdF, void* dFc, void* mom, void* vec, void* vmax, int n, long long ld,
          const double* sqnorm, double alpha, double beta1, double beta2, double eps, int timestep,
          hipStream_t s);

// Accumulate `src` user AoS gradient (float or double; host data must be
// staged to device first) into carry planes: carry[off+c] += src (permuted).
void accumulate_carry(Prec pr, const void* src, int src_f64, void* carry, int plane_off, int nc, const int* perm,
                      int n, long long ld, hipStream_t s);

}  // namespace dmpm
