// C-ABI driver over the reference's own DiffMPMLib3D objects.
//
// Test infrastructure only: this file is compiled (by oracle/Makefile, target
// `ref`) against the reference sources where they lie, into
// oracle/_ref/libdmpm_ref.so.  It contains no reference code; it only builds
// the reference's data structures from flat arrays, calls the reference's
// functions, and copies results back out.  It is used to write golden vectors
// (tests/golden/make_golden.py), to check the oracle's restatement, and as the
// timed CPU baseline.  Nothing in the shipped library links it.
//
// Array conventions (all float32, row-major):
//   x, v           (N,3)
//   F, C, dFc, ... (N,3,3), element (r,c) at [9*p + 3*r + c]
//   grid arrays    node (i,j,k) at i*ny*nz + j*nz + k  (Grid.h:89-94)
#include "pch.h"
#include "PointCloud.h"
#include "Grid.h"
#include "Elasticity.h"
#include "ForwardSimulation.h"
#include "BackPropagation.h"
#include "CompGraph.h"
#include "qrsvd/ImplicitQRSVD.h"
#include "GeometryLoading.h"

#include <omp.h>
#include <cstring>
#include <memory>

using namespace DiffMPMLib3D;

namespace {

Mat3 load_m3(const float* a) {
    Mat3 M;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M(r, c) = a[3 * r + c];
    return M;
}
void store_m3(const Mat3& M, float* a) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[3 * r + c] = M(r, c);
}
Vec3 load_v3(const float* a) { return Vec3(a[0], a[1], a[2]); }
void store_v3(const Vec3& v, float* a) { a[0] = v[0]; a[1] = v[1]; a[2] = v[2]; }

struct RefParticles {
    int n;
    const float *x, *v, *F, *C, *dFc, *m, *vol, *lam, *mu;
};

std::shared_ptr<PointCloud> make_pc(const RefParticles& s) {
    auto pc = std::make_shared<PointCloud>();
    pc->points.resize(s.n);
    for (int p = 0; p < s.n; ++p) {
        MaterialPoint& mp = pc->points[p];
        mp.x = load_v3(s.x + 3 * p);
        mp.v = load_v3(s.v + 3 * p);
        mp.F = load_m3(s.F + 9 * p);
        mp.C = load_m3(s.C + 9 * p);
        mp.dFc = s.dFc ? load_m3(s.dFc + 9 * p) : Mat3::Zero();
        mp.m = s.m[p];
        mp.vol = s.vol[p];
        mp.lam = s.lam[p];
        mp.mu = s.mu[p];
    }
    return pc;
}

std::shared_ptr<Grid> make_grid(const int* dims, float dx, const float* mn) {
    return std::make_shared<Grid>(dims[0], dims[1], dims[2], dx, load_v3(mn));
}

}  // namespace

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }
int ref_max_threads() { return omp_get_max_threads(); }

// ForwardTimeStep (ForwardSimulation.cpp:159-185 in the 8-argument form)
// from one layer: writes layer n+1's x, v, F, C, the P computed for layer n,
// and the grid's m, v after the step.
void ref_forward_step(int n, const float* x, const float* v, const float* F,
                      const float* C, const float* dFc, const float* m,
                      const float* vol, const float* lam, const float* mu,
                      const int* dims, float dx, const float* grid_min,
                      float smoothing, float dt, float drag, const float* f_ext,
                      float* out_x, float* out_v, float* out_F, float* out_C,
                      float* out_P, float* grid_m, float* grid_v) {
    RefParticles s{n, x, v, F, C, dFc, m, vol, lam, mu};
    auto curr = make_pc(s);
    auto next = std::make_shared<PointCloud>(*curr);
    auto grid = make_grid(dims, dx, grid_min);
    SingleThreadMPM::ForwardTimeStep(*next, *curr, *grid, smoothing, dt, drag,
                                     load_v3(f_ext), 0);
    for (int p = 0; p < n; ++p) {
        const MaterialPoint& q = next->points[p];
        store_v3(q.x, out_x + 3 * p);
        store_v3(q.v, out_v + 3 * p);
        store_m3(q.F, out_F + 9 * p);
        store_m3(q.C, out_C + 9 * p);
        if (out_P) store_m3(curr->points[p].P, out_P + 9 * p);
    }
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    for (size_t i = 0; i < G; ++i) {
        if (grid_m) grid_m[i] = grid->nodes[i].m;
        if (grid_v) store_v3(grid->nodes[i].v, grid_v + 3 * i);
    }
}

// One reference adjoint step (BackPropagation.cpp:9-144).  Layer n is given
// by its state; the forward step is run first so that layer n's P, the grid
// and layer n+1's C are exactly what Back_Timestep sees inside CompGraph.
// The adjoint of layer n+1 (dLdx, dLdv, dLdF, dLdC) is an input; layer n's
// adjoint is the output.  Gradients are reset first, as ComputeBackwardPass
// does (CompGraph.cpp:202-205).
void ref_backward_step(int n, const float* x, const float* v, const float* F,
                       const float* C, const float* dFc, const float* m,
                       const float* vol, const float* lam, const float* mu,
                       const int* dims, float dx, const float* grid_min,
                       float smoothing, float dt, float drag, const float* f_ext,
                       const float* gx1, const float* gv1, const float* gF1,
                       const float* gC1, float* gx, float* gv, float* gF,
                       float* gC, float* gP) {
    RefParticles s{n, x, v, F, C, dFc, m, vol, lam, mu};
    CompGraphLayer ln, ln1;
    ln.point_cloud = make_pc(s);
    ln1.point_cloud = std::make_shared<PointCloud>(*ln.point_cloud);
    ln.grid = make_grid(dims, dx, grid_min);
    ln1.grid = std::make_shared<Grid>(*ln.grid);
    SingleThreadMPM::ForwardTimeStep(*ln1.point_cloud, *ln.point_cloud, *ln.grid,
                                     smoothing, dt, drag, load_v3(f_ext), 0);
    ln.grid->ResetGradients();
    ln.point_cloud->ResetGradients();
    for (int p = 0; p < n; ++p) {
        MaterialPoint& q = ln1.point_cloud->points[p];
        q.dLdx = load_v3(gx1 + 3 * p);
        q.dLdv = load_v3(gv1 + 3 * p);
        q.dLdF = load_m3(gF1 + 9 * p);
        q.dLdC = load_m3(gC1 + 9 * p);
    }
    Back_Timestep(ln1, ln, drag, dt, smoothing);
    for (int p = 0; p < n; ++p) {
        const MaterialPoint& q = ln.point_cloud->points[p];
        store_v3(q.dLdx, gx + 3 * p);
        store_v3(q.dLdv, gv + 3 * p);
        store_m3(q.dLdF, gF + 9 * p);
        store_m3(q.dLdC, gC + 9 * p);
        if (gP) store_m3(q.dLdP, gP + 9 * p);
    }
}

// Multi-layer episode the way CompGraph runs it: SetUpCompGraph copies layer 0
// into every layer (CompGraph.cpp:30-40), the forward loop is
// ComputeForwardPass (:131-141), the loss is EndLayerMassLoss (:42-129),
// optional render gradients are added to the last layer (:148-197), and the
// backward loop is ComputeBackwardPass (:199-209).  dt/drag/smoothing/f_ext
// are passed explicitly because CompGraph keeps them private.
// Outputs: final layer x, v, F, C; loss; layer `control_layer`'s dLdx, dLdv,
// dLdF, dLdC.
float ref_episode(int n, const float* x, const float* v, const float* F,
                  const float* C, const float* dFc, const float* m,
                  const float* vol, const float* lam, const float* mu,
                  const int* dims, float dx, const float* grid_min,
                  const float* target_m, int num_layers, int control_layer,
                  float smoothing, float dt, float drag, const float* f_ext,
                  const float* render_gF, const float* render_gx,
                  float* out_x, float* out_v, float* out_F, float* out_C,
                  float* gx, float* gv, float* gF, float* gC) {
    RefParticles s{n, x, v, F, C, dFc, m, vol, lam, mu};
    auto pc0 = make_pc(s);
    auto grid0 = make_grid(dims, dx, grid_min);
    auto tgt = make_grid(dims, dx, grid_min);
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    for (size_t i = 0; i < G; ++i) tgt->nodes[i].m = target_m[i];
    CompGraph cg(pc0, grid0, tgt);
    cg.SetUpCompGraph(num_layers);
    for (int i = 0; i + 1 < num_layers; ++i)
        SingleThreadMPM::ForwardTimeStep(*cg.layers[i + 1].point_cloud,
                                         *cg.layers[i].point_cloud,
                                         *cg.layers[i].grid, smoothing, dt,
                                         drag, load_v3(f_ext), 0);
    float loss = cg.EndLayerMassLoss();
    if (render_gF && render_gx) {
        PointCloud& last = *cg.layers.back().point_cloud;
        for (int p = 0; p < n; ++p) {
            last.points[p].dLdF += load_m3(render_gF + 9 * p);
            last.points[p].dLdx += load_v3(render_gx + 3 * p);
        }
    }
    for (int i = num_layers - 2; i >= control_layer; --i) {
        cg.layers[i].grid->ResetGradients();
        cg.layers[i].point_cloud->ResetGradients();
        Back_Timestep(cg.layers[i + 1], cg.layers[i], drag, dt, smoothing);
    }
    const PointCloud& last = *cg.layers.back().point_cloud;
    const PointCloud& ctl = *cg.layers[control_layer].point_cloud;
    for (int p = 0; p < n; ++p) {
        store_v3(last.points[p].x, out_x + 3 * p);
        store_v3(last.points[p].v, out_v + 3 * p);
        store_m3(last.points[p].F, out_F + 9 * p);
        store_m3(last.points[p].C, out_C + 9 * p);
        store_v3(ctl.points[p].dLdx, gx + 3 * p);
        store_v3(ctl.points[p].dLdv, gv + 3 * p);
        store_m3(ctl.points[p].dLdF, gF + 9 * p);
        store_m3(ctl.points[p].dLdC, gC + 9 * p);
    }
    return loss;
}

// Timed baseline: `steps` ForwardTimeStep calls, then `steps` Back_Timestep
// calls (seeded with dL/dx of the last layer), on the reference's own layers,
// with OpenMP threads as set by ref_set_threads.  Only the two loops are
// timed (omp_get_wtime); fwd_s and bwd_s receive their durations.
void ref_time_steps(int n, const float* x, const float* v, const float* F,
                    const float* C, const float* m, const float* vol,
                    const float* lam, const float* mu, const int* dims,
                    float dx, const float* grid_min, int steps,
                    float smoothing, float dt, float drag, const float* f_ext,
                    const float* seed_gx, double* fwd_s, double* bwd_s) {
    RefParticles s{n, x, v, F, C, nullptr, m, vol, lam, mu};
    std::vector<CompGraphLayer> layers(steps + 1);
    layers[0].point_cloud = make_pc(s);
    layers[0].grid = make_grid(dims, dx, grid_min);
    for (int i = 1; i <= steps; ++i) {
        layers[i].point_cloud = std::make_shared<PointCloud>(*layers[0].point_cloud);
        layers[i].grid = std::make_shared<Grid>(*layers[0].grid);
    }
    double t0 = omp_get_wtime();
    for (int i = 0; i < steps; ++i)
        SingleThreadMPM::ForwardTimeStep(*layers[i + 1].point_cloud, *layers[i].point_cloud,
                                         *layers[i].grid, smoothing, dt, drag, load_v3(f_ext), 0);
    double t1 = omp_get_wtime();
    PointCloud& last = *layers[steps].point_cloud;
    last.ResetGradients();
    for (int p = 0; p < n; ++p) last.points[p].dLdx = load_v3(seed_gx + 3 * p);
    double t2 = omp_get_wtime();
    for (int i = steps - 1; i >= 0; --i) {
        layers[i].grid->ResetGradients();
        layers[i].point_cloud->ResetGradients();
        Back_Timestep(layers[i + 1], layers[i], drag, dt, smoothing);
    }
    double t3 = omp_get_wtime();
    *fwd_s = t1 - t0;
    *bwd_s = t3 - t2;
}

// EndLayerMassLoss (CompGraph.cpp:42-129) on a single layer: returns the loss,
// writes the particles' dLdx and the grid's per-node mass and dLdm.
float ref_mass_loss(int n, const float* x, const float* v, const float* F,
                    const float* C, const float* m, const float* vol,
                    const float* lam, const float* mu, const int* dims,
                    float dx, const float* grid_min, const float* target_m,
                    float* gx, float* grid_m, float* grid_dLdm) {
    RefParticles s{n, x, v, F, C, nullptr, m, vol, lam, mu};
    auto pc = make_pc(s);
    auto grid = make_grid(dims, dx, grid_min);
    auto tgt = make_grid(dims, dx, grid_min);
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    for (size_t i = 0; i < G; ++i) tgt->nodes[i].m = target_m[i];
    CompGraph cg(pc, grid, tgt);
    float loss = cg.EndLayerMassLoss();
    const PointCloud& out = *cg.layers.back().point_cloud;
    const Grid& g = *cg.layers.back().grid;
    for (int p = 0; p < n; ++p) store_v3(out.points[p].dLdx, gx + 3 * p);
    for (size_t i = 0; i < G; ++i) {
        if (grid_m) grid_m[i] = g.nodes[i].m;
        if (grid_dLdm) grid_dLdm[i] = g.nodes[i].dLdm;
    }
    return loss;
}

// P2G with dt = 0, drag = 0 into a fresh grid (bind.cpp:546-549): node masses.
void ref_p2g_mass(int n, const float* x, const float* m, const int* dims,
                  float dx, const float* grid_min, float* grid_m) {
    auto pc = std::make_shared<PointCloud>();
    pc->points.resize(n);
    for (int p = 0; p < n; ++p) {
        pc->points[p].x = load_v3(x + 3 * p);
        pc->points[p].m = m[p];
    }
    auto grid = make_grid(dims, dx, grid_min);
    SingleThreadMPM::P2G(*pc, *grid, 0.f, 0.f);
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    for (size_t i = 0; i < G; ++i) grid_m[i] = grid->nodes[i].m;
}

// CalculatePointCloudVolumes (ForwardSimulation.cpp:234-271).
void ref_volumes(int n, const float* x, const float* m, const int* dims,
                 float dx, const float* grid_min, float* vol, float* grid_m) {
    auto pc = std::make_shared<PointCloud>();
    pc->points.resize(n);
    for (int p = 0; p < n; ++p) {
        pc->points[p].x = load_v3(x + 3 * p);
        pc->points[p].m = m[p];
    }
    auto grid = make_grid(dims, dx, grid_min);
    SingleThreadMPM::CalculatePointCloudVolumes(*pc, *grid);
    for (int p = 0; p < n; ++p) vol[p] = pc->points[p].vol;
    if (grid_m) {
        const size_t G = (size_t)dims[0] * dims[1] * dims[2];
        for (size_t i = 0; i < G; ++i) grid_m[i] = grid->nodes[i].m;
    }
}

// Descend_Adam (PointCloud.cpp:83-147) in place on dFc and the three moment
// arrays, given dLdF.
void ref_adam(int n, const float* dLdF, float* dFc, float* mom, float* vec,
              float* vmax, float alpha, float grad_norm, float beta1,
              float beta2, float eps, int timestep) {
    PointCloud pc;
    pc.points.resize(n);
    for (int p = 0; p < n; ++p) {
        MaterialPoint& q = pc.points[p];
        q.dLdF = load_m3(dLdF + 9 * p);
        q.dFc = load_m3(dFc + 9 * p);
        q.momentum = load_m3(mom + 9 * p);
        q.vector = load_m3(vec + 9 * p);
        q.vector_max = load_m3(vmax + 9 * p);
    }
    pc.Descend_Adam(alpha, grad_norm, beta1, beta2, eps, timestep);
    for (int p = 0; p < n; ++p) {
        const MaterialPoint& q = pc.points[p];
        store_m3(q.dFc, dFc + 9 * p);
        store_m3(q.momentum, mom + 9 * p);
        store_m3(q.vector, vec + 9 * p);
        store_m3(q.vector_max, vmax + 9 * p);
    }
}

// Global dLdF norm (PointCloud.cpp:24-35).
float ref_dLdF_norm(int n, const float* dLdF) {
    PointCloud pc;
    pc.points.resize(n);
    for (int p = 0; p < n; ++p) pc.points[p].dLdF = load_m3(dLdF + 9 * p);
    return pc.Compute_dLdF_Norm();
}

// Elasticity primitives (Elasticity.cpp).
void ref_pk_fcr(int n, const float* F, const float* lam, const float* mu,
                float* P) {
    for (int p = 0; p < n; ++p)
        store_m3(PK_FixedCorotatedElasticity(load_m3(F + 9 * p), lam[p], mu[p]),
                 P + 9 * p);
}
void ref_pk_fcr_hvp(int n, const float* F, const float* lam, const float* mu,
                    const float* dF, float* out) {
    for (int p = 0; p < n; ++p)
        store_m3(d2_FCE_psi_dF2_mult_by_dF(load_m3(F + 9 * p), lam[p], mu[p],
                                           load_m3(dF + 9 * p)),
                 out + 9 * p);
}
void ref_energy_fcr(int n, const float* F, const float* lam, const float* mu,
                    float* psi) {
    for (int p = 0; p < n; ++p)
        psi[p] = FixedCorotatedElasticity(load_m3(F + 9 * p), lam[p], mu[p]);
}
void ref_polar(int n, const float* F, float* R, float* S) {
    for (int p = 0; p < n; ++p) {
        Mat3 r, s;
        JIXIE::polarDecomposition(load_m3(F + 9 * p), r, s);
        store_m3(r, R + 9 * p);
        store_m3(s, S + 9 * p);
    }
}
// PointCloud::WriteEntirePointCloudToBinaryFile (PointCloud.cpp:382-394) of
// a point cloud holding exactly the given fields (MaterialPoint.h:55-71):
// vectors (N,3), matrices (N,3,3), scalars (N,).  Returns 0 on success.
int ref_write_cereal(int n, const float* x, const float* v, const float* F, const float* C, const float* P,
                     const float* m, const float* vol, const float* lam, const float* mu, const float* dFc,
                     const float* dLdx, const float* dLdv, const float* dLdF, const float* mom,
                     const float* vec, const float* vmax, const char* path) {
    PointCloud pc;
    pc.points.resize(n);
    for (int p = 0; p < n; ++p) {
        MaterialPoint& mp = pc.points[p];
        mp.x = load_v3(x + 3 * p);
        mp.v = load_v3(v + 3 * p);
        mp.F = load_m3(F + 9 * p);
        mp.C = load_m3(C + 9 * p);
        mp.P = load_m3(P + 9 * p);
        mp.m = m[p];
        mp.vol = vol[p];
        mp.lam = lam[p];
        mp.mu = mu[p];
        mp.dFc = load_m3(dFc + 9 * p);
        mp.dLdx = load_v3(dLdx + 3 * p);
        mp.dLdv = load_v3(dLdv + 3 * p);
        mp.dLdF = load_m3(dLdF + 9 * p);
        mp.momentum = load_m3(mom + 9 * p);
        mp.vector = load_m3(vec + 9 * p);
        mp.vector_max = load_m3(vmax + 9 * p);
    }
    pc.WriteEntirePointCloudToBinaryFile(path);
    return 0;
}

// PointCloud::ReadEntirePointCloudFromBinaryFile (PointCloud.cpp:396-410):
// the number of points of the file, then (second call, arrays non-null) its
// x, F, P, m and vector_max fields.
int ref_read_cereal(const char* path, int n_max, float* x, float* F, float* P, float* m, float* vmax) {
    PointCloud pc;
    if (!pc.ReadEntirePointCloudFromBinaryFile(path)) return -1;
    const int n = (int)pc.points.size();
    if (!x) return n;
    for (int p = 0; p < n && p < n_max; ++p) {
        const MaterialPoint& mp = pc.points[p];
        store_v3(mp.x, x + 3 * p);
        store_m3(mp.F, F + 9 * p);
        store_m3(mp.P, P + 9 * p);
        m[p] = mp.m;
        store_m3(mp.vector_max, vmax + 9 * p);
    }
    return n;
}

// GeometryLoading::LoadMPMPointCloudFromObj (GeometryLoading.cpp:85-143): the
// number of points it generates from the OBJ mesh, and (arrays non-null, at
// most n_max) their positions, masses and volumes.  -1 if it fails.
int ref_load_obj(const char* path, float point_dx, float density, float lam, float mu, int n_max, float* x,
                 float* m, float* vol) {
    std::shared_ptr<PointCloud> pc;
    if (!GeometryLoading::LoadMPMPointCloudFromObj(path, pc, point_dx, density, lam, mu)) return -1;
    const int n = (int)pc->points.size();
    if (!x) return n;
    for (int p = 0; p < n && p < n_max; ++p) {
        store_v3(pc->points[p].x, x + 3 * p);
        m[p] = pc->points[p].m;
        vol[p] = pc->points[p].vol;
    }
    return n;
}

void ref_lame(float E, float nu, float* lam, float* mu) {
    CalculateLameParameters(E, nu, *lam, *mu);
}

}  // extern "C"

extern "C" {

// `steps` ForwardTimeStep calls from one layer (ForwardSimulation.cpp:159-185,
// each from a layer whose dFc is the given one for the first step and zero
// after, as P_op_2 writes it, :156), then -- when a seed is given -- the
// seed as the last layer's adjoint and `steps` Back_Timestep calls
// (BackPropagation.cpp:9-144) with the per-layer gradient reset of
// ComputeBackwardPass (CompGraph.cpp:199-209).  Outputs: the last layer's
// x, v, F, C and (with a seed) layer 0's dLdx, dLdv, dLdF, dLdC.
void ref_rollout(int n, const float* x, const float* v, const float* F, const float* C, const float* dFc,
                 const float* m, const float* vol, const float* lam, const float* mu, const int* dims, float dx,
                 const float* grid_min, int steps, float smoothing, float dt, float drag, const float* f_ext,
                 const float* sx, const float* sv, const float* sF, const float* sC, float* out_x, float* out_v,
                 float* out_F, float* out_C, float* gx, float* gv, float* gF, float* gC) {
    RefParticles s{n, x, v, F, C, dFc, m, vol, lam, mu};
    std::vector<CompGraphLayer> layers(steps + 1);
    layers[0].point_cloud = make_pc(s);
    layers[0].grid = make_grid(dims, dx, grid_min);
    for (int i = 1; i <= steps; ++i) {
        layers[i].point_cloud = std::make_shared<PointCloud>(*layers[0].point_cloud);
        layers[i].grid = std::make_shared<Grid>(*layers[0].grid);
    }
    for (int i = 0; i < steps; ++i)
        SingleThreadMPM::ForwardTimeStep(*layers[i + 1].point_cloud, *layers[i].point_cloud, *layers[i].grid,
                                         smoothing, dt, drag, load_v3(f_ext), 0);
    const PointCloud& last = *layers[steps].point_cloud;
    for (int p = 0; p < n; ++p) {
        store_v3(last.points[p].x, out_x + 3 * p);
        store_v3(last.points[p].v, out_v + 3 * p);
        store_m3(last.points[p].F, out_F + 9 * p);
        store_m3(last.points[p].C, out_C + 9 * p);
    }
    if (!sx) return;
    PointCloud& lp = *layers[steps].point_cloud;
    lp.ResetGradients();
    for (int p = 0; p < n; ++p) {
        lp.points[p].dLdx = load_v3(sx + 3 * p);
        lp.points[p].dLdv = load_v3(sv + 3 * p);
        lp.points[p].dLdF = load_m3(sF + 9 * p);
        lp.points[p].dLdC = load_m3(sC + 9 * p);
    }
    for (int i = steps - 1; i >= 0; --i) {
        layers[i].grid->ResetGradients();
        layers[i].point_cloud->ResetGradients();
        Back_Timestep(layers[i + 1], layers[i], drag, dt, smoothing);
    }
    const PointCloud& l0 = *layers[0].point_cloud;
    for (int p = 0; p < n; ++p) {
        store_v3(l0.points[p].dLdx, gx + 3 * p);
        store_v3(l0.points[p].dLdv, gv + 3 * p);
        store_m3(l0.points[p].dLdF, gF + 9 * p);
        store_m3(l0.points[p].dLdC, gC + 9 * p);
    }
}

// CompGraph::OptimizeDefGradControlSequence (CompGraph.cpp:213-368; bound as
// run_optimization, bind.cpp:165-172) on CompGraph(pc, grid, target), then,
// when st_iters > 0, CompGraph::OptimizeSingleTimestep(st_t, st_iters, 0,
// st_alpha, st_ls) (CompGraph.cpp:369-465) on the same graph, whose
// simulation parameters the first call set.  Outputs: the end-layer mass
// loss after each call (out_loss[0], out_loss[1]); every layer's dFc after
// the first call (num_steps x n x 9) and after the second (same layout);
// the Adam state (momentum, vector, vector_max) of layer st_t after the
// second call (3 x n x 9).
void ref_optimize(int n, const float* x, const float* v, const float* F, const float* C, const float* m,
                  const float* vol, const float* lam, const float* mu, const int* dims, float dx,
                  const float* grid_min, const float* target_m, int num_steps, float dt, float drag,
                  const float* f_ext, int control_stride, int max_gd_iters, int max_ls_iters, float initial_alpha,
                  float gd_tol, float smoothing, int st_t, int st_iters, float st_alpha, int st_ls, float* out_loss,
                  float* dfc_run, float* dfc_single, float* adam_single) {
    RefParticles s{n, x, v, F, C, nullptr, m, vol, lam, mu};
    auto pc0 = make_pc(s);
    auto grid0 = make_grid(dims, dx, grid_min);
    auto tgt = make_grid(dims, dx, grid_min);
    const size_t G = (size_t)dims[0] * dims[1] * dims[2];
    for (size_t i = 0; i < G; ++i) tgt->nodes[i].m = target_m[i];
    CompGraph cg(pc0, grid0, tgt);
    cg.OptimizeDefGradControlSequence(num_steps, dt, drag, load_v3(f_ext), control_stride, max_gd_iters,
                                      max_ls_iters, initial_alpha, gd_tol, smoothing, 0);
    auto dump_dfc = [&](float* out) {
        for (int l = 0; l < num_steps; ++l)
            for (int p = 0; p < n; ++p)
                store_m3(cg.layers[l].point_cloud->points[p].dFc, out + ((size_t)l * n + p) * 9);
    };
    out_loss[0] = cg.EndLayerMassLoss();
    dump_dfc(dfc_run);
    if (st_iters <= 0) return;
    cg.OptimizeSingleTimestep(st_t, st_iters, 0, st_alpha, st_ls);
    out_loss[1] = cg.EndLayerMassLoss();
    dump_dfc(dfc_single);
    const PointCloud& pc = *cg.layers[st_t].point_cloud;
    for (int p = 0; p < n; ++p) {
        store_m3(pc.points[p].momentum, adam_single + ((size_t)0 * n + p) * 9);
        store_m3(pc.points[p].vector, adam_single + ((size_t)1 * n + p) * 9);
        store_m3(pc.points[p].vector_max, adam_single + ((size_t)2 * n + p) * 9);
    }
}

}  // extern "C"
