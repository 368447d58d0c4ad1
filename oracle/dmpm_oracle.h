/* CPU oracle for the differentiable MPM step.  TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch restatement of the reference's DiffMPMLib3D step and its
 * adjoint, used as the checker for the HIP library: only tests/, the smoke
 * check and bench.py's CPU-baseline leg load it.  The product path never
 * calls it.
 *
 * Every function comes in an _f32 and an _f64 flavour; the arrays passed must
 * be of that scalar type.  Array conventions (row-major):
 *   x, v           (N,3)
 *   F, C, dFc, ... (N,3,3), element (r,c) at [9*p + 3*r + c]
 *   grid arrays    node (i,j,k) at i*ny*nz + j*nz + k   (Grid.h:89-94)
 */
#ifndef DMPM_ORACLE_H
#define DMPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dims[3];        /* node counts per axis */
    double dx;
    double grid_min[3];
    double dt;
    double drag;
    double smoothing;   /* F' = (1-s)(I+dtC)Ft + s F   (ForwardSimulation.cpp:131-133) */
    double zeta;        /* bulk viscosity coefficient   (ForwardSimulation.cpp:13) */
    double jmin;        /* trust-region J_min           (ForwardSimulation.cpp:13) */
    double f_ext[3];
    int order;          /* 3 = cubic B-spline (reference), 2 = quadratic */
    int model;          /* 0 = fixed-corotated (reference), 1 = Neo-Hookean */
    int trust_region;   /* 1 = J trust region on (reference) */
    int wall_cells;     /* sticky walls: v=0 on nodes within this many cells of a face; 0 = off */
    int adjoint;        /* 0 = exact VJP, 1 = reference formulas (BackPropagation.cpp) */
} orc_params;

#define ORC_DECL(S, T)                                                          \
    void orc_forward_step_##S(const orc_params* prm, int n, const T* x,         \
        const T* v, const T* F, const T* C, const T* dFc, const T* m,          \
        const T* vol, const T* lam, const T* mu, T* out_x, T* out_v,            \
        T* out_F, T* out_C, T* out_P, T* grid_m, T* grid_v);                    \
    void orc_backward_step_##S(const orc_params* prm, int n, const T* x,        \
        const T* v, const T* F, const T* C, const T* dFc, const T* m,          \
        const T* vol, const T* lam, const T* mu, const T* gx1, const T* gv1,   \
        const T* gF1, const T* gC1, T* gx, T* gv, T* gF, T* gC, T* g_dFc,      \
        T* gP);                                                                 \
    void orc_pk1_##S(int model, int n, const T* F, const T* lam, const T* mu,   \
        T* P);                                                                  \
    void orc_pk1_hvp_##S(int model, int n, const T* F, const T* lam,            \
        const T* mu, const T* dF, T* out);                                      \
    void orc_energy_##S(int model, int n, const T* F, const T* lam,             \
        const T* mu, T* psi);                                                   \
    void orc_svd_##S(int n, const T* F, T* U, T* sigma, T* V);                  \
    void orc_polar_##S(int n, const T* F, T* R, T* S);                          \
    void orc_p2g_mass_##S(const orc_params* prm, int n, const T* x,             \
        const T* m, T* grid_m);                                                 \
    void orc_volumes_##S(const orc_params* prm, int n, const T* x,              \
        const T* m, T* vol, T* grid_m);                                         \
    double orc_mass_loss_##S(const orc_params* prm, int n, const T* x,          \
        const T* m, const T* target_m, T* gx, T* grid_m, T* grid_dLdm);         \
    void orc_adam_##S(int n, const T* dLdF, T* dFc, T* mom, T* vec, T* vmax,    \
        double alpha, double grad_norm, double beta1, double beta2,            \
        double eps, int timestep);                                              \
    double orc_dLdF_norm_##S(int n, const T* dLdF);                             \
    void orc_bspline_##S(int order, int n, const T* u, T* w, T* dw);

ORC_DECL(f32, float)
ORC_DECL(f64, double)

#undef ORC_DECL

#ifdef __cplusplus
}
#endif
#endif
