// CPU oracle: a restatement of the reference's differentiable MPM step.
// TEST INFRASTRUCTURE ONLY (see dmpm_oracle.h).  Serial in every scatter so
// results are deterministic; gathers are independent per particle.
//
// Each function cites the reference file:line whose behaviour it restates.
// Two knobs select the reference's own form: order = 3 (cubic B-spline,
// Interpolation.h:5-30), model = 0 (fixed-corotated, Elasticity.cpp:45-55),
// trust_region = 1 and zeta = 0.25 (ForwardSimulation.cpp:13,136-152),
// adjoint = 1 (BackPropagation.cpp:9-144 verbatim, quirks included).  The
// other settings are the "plain" method of the configs: quadratic B-spline,
// Neo-Hookean, no trust region, sticky walls, exact adjoint.
#include "dmpm_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

template <typename T>
struct M3 {
    T a[9];
    T& operator()(int r, int c) { return a[3 * r + c]; }
    T operator()(int r, int c) const { return a[3 * r + c]; }
};
template <typename T>
struct V3 {
    T a[3];
    T& operator[](int i) { return a[i]; }
    T operator[](int i) const { return a[i]; }
};

template <typename T> M3<T> zero3() { M3<T> M; for (auto& e : M.a) e = T(0); return M; }
template <typename T> M3<T> eye3() { M3<T> M = zero3<T>(); M(0,0) = M(1,1) = M(2,2) = T(1); return M; }
template <typename T> M3<T> ld(const T* p) { M3<T> M; std::memcpy(M.a, p, sizeof(M.a)); return M; }
template <typename T> V3<T> ldv(const T* p) { V3<T> v; std::memcpy(v.a, p, sizeof(v.a)); return v; }
template <typename T> void st(const M3<T>& M, T* p) { std::memcpy(p, M.a, sizeof(M.a)); }
template <typename T> void stv(const V3<T>& v, T* p) { std::memcpy(p, v.a, sizeof(v.a)); }

template <typename T> M3<T> add(const M3<T>& A, const M3<T>& B) { M3<T> R; for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] + B.a[i]; return R; }
template <typename T> M3<T> sub(const M3<T>& A, const M3<T>& B) { M3<T> R; for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] - B.a[i]; return R; }
template <typename T> M3<T> scl(T s, const M3<T>& A) { M3<T> R; for (int i = 0; i < 9; ++i) R.a[i] = s * A.a[i]; return R; }
template <typename T> M3<T> tr(const M3<T>& A) { M3<T> R; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) R(r,c) = A(c,r); return R; }
template <typename T> M3<T> mul(const M3<T>& A, const M3<T>& B) {
    M3<T> R;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            R(r, c) = A(r, 0) * B(0, c) + A(r, 1) * B(1, c) + A(r, 2) * B(2, c);
    return R;
}
template <typename T> V3<T> mulv(const M3<T>& A, const V3<T>& v) {
    V3<T> r;
    for (int i = 0; i < 3; ++i) r[i] = A(i, 0) * v[0] + A(i, 1) * v[1] + A(i, 2) * v[2];
    return r;
}
template <typename T> T ddot(const M3<T>& A, const M3<T>& B) { T s = 0; for (int i = 0; i < 9; ++i) s += A.a[i] * B.a[i]; return s; }
template <typename T> T dotv(const V3<T>& a, const V3<T>& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <typename T> T trace(const M3<T>& A) { return A(0,0) + A(1,1) + A(2,2); }
template <typename T> T fnorm(const M3<T>& A) { return std::sqrt(ddot(A, A)); }
template <typename T> M3<T> outer(const V3<T>& a, const V3<T>& b) { M3<T> R; for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) R(r,c) = a[r] * b[c]; return R; }

// 3x3 determinant by cofactor expansion along the first row.
template <typename T> T det3(const M3<T>& m) {
    return m(0,0) * (m(1,1) * m(2,2) - m(1,2) * m(2,1))
         - m(0,1) * (m(1,0) * m(2,2) - m(1,2) * m(2,0))
         + m(0,2) * (m(1,0) * m(2,1) - m(1,1) * m(2,0));
}
// Cofactor matrix: cof(A) = det(A) A^{-T}.
template <typename T> M3<T> cof3(const M3<T>& m) {
    M3<T> c;
    c(0,0) = m(1,1) * m(2,2) - m(1,2) * m(2,1);
    c(0,1) = m(1,2) * m(2,0) - m(1,0) * m(2,2);
    c(0,2) = m(1,0) * m(2,1) - m(1,1) * m(2,0);
    c(1,0) = m(0,2) * m(2,1) - m(0,1) * m(2,2);
    c(1,1) = m(0,0) * m(2,2) - m(0,2) * m(2,0);
    c(1,2) = m(0,1) * m(2,0) - m(0,0) * m(2,1);
    c(2,0) = m(0,1) * m(1,2) - m(0,2) * m(1,1);
    c(2,1) = m(0,2) * m(1,0) - m(0,0) * m(1,2);
    c(2,2) = m(0,0) * m(1,1) - m(0,1) * m(1,0);
    return c;
}
template <typename T> M3<T> inv_tr(const M3<T>& m) { return scl(T(1) / det3(m), cof3(m)); }

template <typename T, typename U> M3<U> cast(const M3<T>& A) { M3<U> R; for (int i = 0; i < 9; ++i) R.a[i] = (U)A.a[i]; return R; }

// ---------------------------------------------------------------------------
// 3x3 SVD by one-sided Jacobi (Hestenes) in double, iterated to convergence.
// Convention of the reference's polar decomposition (ImplicitQRSVD.h:675-680,
// 852-875): U and V are rotations, singular values sorted by decreasing
// magnitude, and only the last one may be negative.
void svd3(const M3<double>& F, M3<double>& U, double sig[3], M3<double>& V) {
    M3<double> A = F;
    V = eye3<double>();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, tot = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < 3; ++i) {
                    al += A(i, p) * A(i, p);
                    be += A(i, q) * A(i, q);
                    ga += A(i, p) * A(i, q);
                }
                tot += al * be;
                off += ga * ga;
                if (ga == 0.0) continue;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < 3; ++i) {
                    double ap = A(i, p), aq = A(i, q);
                    A(i, p) = c * ap - s * aq;
                    A(i, q) = s * ap + c * aq;
                    double vp = V(i, p), vq = V(i, q);
                    V(i, p) = c * vp - s * vq;
                    V(i, q) = s * vp + c * vq;
                }
            }
        if (off <= 1e-34 * tot || tot == 0.0) break;
    }
    double s[3];
    for (int j = 0; j < 3; ++j) s[j] = std::sqrt(A(0, j) * A(0, j) + A(1, j) * A(1, j) + A(2, j) * A(2, j));
    // sort by decreasing singular value
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int a, int b) { return s[a] > s[b]; });
    M3<double> A2, V2;
    double s2[3];
    for (int j = 0; j < 3; ++j) {
        s2[j] = s[idx[j]];
        for (int i = 0; i < 3; ++i) { A2(i, j) = A(i, idx[j]); V2(i, j) = V(i, idx[j]); }
    }
    // U columns = A columns / sigma; complete degenerate columns.
    M3<double> Uu = zero3<double>();
    const double tiny = 1e-300;
    for (int j = 0; j < 2; ++j) {
        if (s2[j] > tiny) for (int i = 0; i < 3; ++i) Uu(i, j) = A2(i, j) / s2[j];
    }
    if (!(s2[0] > tiny)) { Uu = eye3<double>(); }
    else if (!(s2[1] > tiny)) {
        // pick any unit vector orthogonal to u0
        V3<double> u0{{Uu(0,0), Uu(1,0), Uu(2,0)}};
        V3<double> e{{1, 0, 0}};
        if (std::fabs(u0[0]) > 0.9) e = V3<double>{{0, 1, 0}};
        double d = dotv(u0, e);
        V3<double> u1{{e[0] - d * u0[0], e[1] - d * u0[1], e[2] - d * u0[2]}};
        double nn = std::sqrt(dotv(u1, u1));
        for (int i = 0; i < 3; ++i) Uu(i, 1) = u1[i] / nn;
    }
    {
        V3<double> u0{{Uu(0,0), Uu(1,0), Uu(2,0)}}, u1{{Uu(0,1), Uu(1,1), Uu(2,1)}};
        V3<double> u2{{u0[1] * u1[2] - u0[2] * u1[1], u0[2] * u1[0] - u0[0] * u1[2], u0[0] * u1[1] - u0[1] * u1[0]}};
        // sigma_2 carries the sign of A2's third column along u2
        double proj = A2(0, 2) * u2[0] + A2(1, 2) * u2[1] + A2(2, 2) * u2[2];
        for (int i = 0; i < 3; ++i) Uu(i, 2) = u2[i];
        s2[2] = proj;
    }
    // make V a rotation; flipping V's column j flips sigma_j's sign
    if (det3(V2) < 0) {
        for (int i = 0; i < 3; ++i) V2(i, 2) = -V2(i, 2);
        s2[2] = -s2[2];
    }
    U = Uu;
    V = V2;
    for (int j = 0; j < 3; ++j) sig[j] = s2[j];
}

template <typename T>
void polar3(const M3<T>& F, M3<T>& R, M3<T>& S) {
    M3<double> U, V;
    double s[3];
    svd3(cast<T, double>(F), U, s, V);
    M3<double> Rd = mul(U, tr(V));
    M3<double> Sd = zero3<double>();
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < 3; ++k) Sd(r, c) += V(r, k) * s[k] * V(c, k);
    R = cast<double, T>(Rd);
    S = cast<double, T>(Sd);
}

// ---------------------------------------------------------------------------
// Constitutive models.
// Fixed-corotated PK1 (Elasticity.cpp:45-55): P = 2mu(F-R) + lam(J-1)J F^{-T},
// J clamped >= 1e-6.
// Neo-Hookean: P = mu(F - F^{-T}) + lam log J F^{-T}, J clamped >= 1e-6 in log.
template <typename T>
M3<T> pk1(int model, const M3<T>& F, T lam, T mu) {
    if (model == 0) {
        M3<T> R, S;
        polar3(F, R, S);
        T J = std::max(det3(F), T(1e-6));
        return add(scl(T(2) * mu, sub(F, R)), scl(lam * (J - T(1)) * J, inv_tr(F)));
    }
    T J = det3(F);
    M3<T> Fit = inv_tr(F);
    T logJ = std::log(std::max(J, T(1e-6)));
    return add(scl(mu, sub(F, Fit)), scl(lam * logJ, Fit));
}

// Hessian action dP = d^2psi/dF^2 : dF.
// FCR: the reference's analytic form (Elasticity.cpp:57-117): d(J F^{-T})/dF
// from d_JFit_dF (:12-31, unclamped J), rotation derivative from
// (tr S I - S) w = axial(R^T dF), regularised when |det| < 1e-6.
// NH: dP = mu dF + lam (F^{-T}:dF) F^{-T} + (mu - lam log J) F^{-T} dF^T F^{-T}.
template <typename T>
M3<T> pk1_hvp(int model, const M3<T>& F, T lam, T mu, const M3<T>& dF) {
    if (model == 0) {
        M3<T> R, S;
        polar3(F, R, S);
        T Jraw = det3(F);
        M3<T> Fit = inv_tr(F);
        T J = std::max(Jraw, T(1e-6));
        M3<T> ret = scl(T(2) * mu, dF);
        ret = add(ret, scl(lam * J * ddot(scl(J, Fit), dF), Fit));
        M3<T> term = scl(Jraw, sub(scl(ddot(Fit, dF), Fit), mul(mul(Fit, tr(dF)), Fit)));
        ret = add(ret, scl(lam * (J - T(1)), term));
        M3<T> A;
        A(0,0) = S(1,1) + S(2,2); A(0,1) = -S(0,1);        A(0,2) = -S(0,2);
        A(1,0) = -S(0,1);        A(1,1) = S(0,0) + S(2,2); A(1,2) = -S(1,2);
        A(2,0) = -S(0,2);        A(2,1) = -S(1,2);        A(2,2) = S(0,0) + S(1,1);
        if (std::fabs(det3(A)) < 1e-6) { A(0,0) += T(1e-4); A(1,1) += T(1e-4); A(2,2) += T(1e-4); }
        M3<T> t = mul(tr(R), dF);
        // solve A w = b in double (any exact solver; the reference uses QR)
        M3<double> Ad = cast<T, double>(A);
        double bd[3] = {(double)t(2,1) - (double)t(1,2), (double)t(0,2) - (double)t(2,0), (double)t(1,0) - (double)t(0,1)};
        M3<double> Ai = tr(inv_tr(Ad));  // A^{-1}
        double w[3];
        for (int i = 0; i < 3; ++i) w[i] = Ai(i,0) * bd[0] + Ai(i,1) * bd[1] + Ai(i,2) * bd[2];
        M3<T> W;
        W(0,0) = 0;          W(0,1) = -(T)w[2];   W(0,2) = (T)w[1];
        W(1,0) = (T)w[2];    W(1,1) = 0;          W(1,2) = -(T)w[0];
        W(2,0) = -(T)w[1];   W(2,1) = (T)w[0];    W(2,2) = 0;
        ret = sub(ret, scl(T(2) * mu, mul(R, W)));
        return ret;
    }
    T J = det3(F);
    M3<T> Fit = inv_tr(F);
    T logJ = std::log(std::max(J, T(1e-6)));
    M3<T> ret = scl(mu, dF);
    ret = add(ret, scl(lam * ddot(Fit, dF), Fit));
    ret = add(ret, scl(mu - lam * logJ, mul(mul(Fit, tr(dF)), Fit)));
    return ret;
}

// Elastic energy: FCR mu sum(sigma-1)^2 + lam/2 (J-1)^2 (Elasticity.cpp:33-43);
// NH mu/2 (|F|^2 - 3) - mu log J + lam/2 log^2 J.
template <typename T>
T energy(int model, const M3<T>& F, T lam, T mu) {
    T J = det3(F);
    if (model == 0) {
        M3<double> U, V;
        double s[3];
        svd3(cast<T, double>(F), U, s, V);
        T ss = 0;
        for (double si : s) ss += (T(si) - T(1)) * (T(si) - T(1));
        return mu * ss + T(0.5) * lam * (J - T(1)) * (J - T(1));
    }
    T logJ = std::log(std::max(J, T(1e-6)));
    return T(0.5) * mu * (ddot(F, F) - T(3)) - mu * logJ + T(0.5) * lam * logJ * logJ;
}

// ---------------------------------------------------------------------------
// B-splines.  Cubic: Interpolation.h:5-30.  Quadratic: N(u) = 3/4 - u^2 for
// |u| < 1/2, (3/2 - |u|)^2 / 2 for 1/2 <= |u| < 3/2.
template <typename T>
T bs(int order, T u) {
    T a = std::fabs(u);
    if (order == 3) {
        if (a < T(1)) return T(0.5) * a * a * a - a * a + T(2) / T(3);
        if (a < T(2)) return (T(2) - a) * (T(2) - a) * (T(2) - a) / T(6);
        return T(0);
    }
    if (a < T(0.5)) return T(0.75) - a * a;
    if (a < T(1.5)) return T(0.5) * (T(1.5) - a) * (T(1.5) - a);
    return T(0);
}
template <typename T>
T dbs(int order, T u) {
    T a = std::fabs(u);
    if (order == 3) {
        if (a < T(1)) return T(1.5) * u * a - T(2) * u;
        if (a < T(2)) return -u * a / T(2) + T(2) * u - T(2) * u / a;
        return T(0);
    }
    if (a < T(0.5)) return T(-2) * u;
    if (a < T(1.5)) return (u > 0 ? T(-1) : T(1)) * (T(1.5) - a);
    return T(0);
}

// Stencil of one particle: base node and per-axis offsets.  Cubic base =
// floor((x-min)/dx) - 1 (ForwardSimulation.cpp:32-34); quadratic base =
// floor((x-min)/dx - 1/2).  D^{-1} = 3/dx^2 (cubic, ForwardSimulation.cpp:61)
// or 4/dx^2 (quadratic).
template <typename T>
struct Stencil {
    int base[3];
    int K;
};

template <typename T>
struct Ctx {
    const orc_params* p;
    int nx, ny, nz;
    T dx, mn[3];
    T dinv;
    int K;
    explicit Ctx(const orc_params* prm) : p(prm) {
        nx = prm->dims[0]; ny = prm->dims[1]; nz = prm->dims[2];
        dx = (T)prm->dx;
        for (int i = 0; i < 3; ++i) mn[i] = (T)prm->grid_min[i];
        K = prm->order == 3 ? 4 : 3;
        dinv = (prm->order == 3 ? T(3) : T(4)) / (dx * dx);
    }
    size_t G() const { return (size_t)nx * ny * nz; }
    void base(const V3<T>& x, int b[3]) const {
        for (int a = 0; a < 3; ++a) {
            T rel = x[a] - mn[a];
            if (p->order == 3) b[a] = (int)std::floor(rel / dx) - 1;
            else b[a] = (int)std::floor(rel / dx - T(0.5));
        }
    }
    bool inb(int i, int j, int k) const { return 0 <= i && i < nx && 0 <= j && j < ny && 0 <= k && k < nz; }
    size_t id(int i, int j, int k) const { return (size_t)i * ny * nz + (size_t)j * nz + k; }
    bool wall(int i, int j, int k) const {
        int w = p->wall_cells;
        if (w <= 0) return false;
        return i < w || j < w || k < w || i >= nx - w || j >= ny - w || k >= nz - w;
    }
    // node position, as Grid.cpp:19 computes it: min + (i,j,k) * dx
    V3<T> xnode(int i, int j, int k) const {
        return V3<T>{{mn[0] + (T)i * dx, mn[1] + (T)j * dx, mn[2] + (T)k * dx}};
    }
    // iterate stencil nodes in the reference's order (i outer, k inner)
    template <typename Fn>
    void for_stencil(const V3<T>& x, bool grad, Fn fn) const {
        int b[3];
        base(x, b);
        for (int a = 0; a < K; ++a)
            for (int bb = 0; bb < K; ++bb)
                for (int c = 0; c < K; ++c) {
                    int i = b[0] + a, j = b[1] + bb, k = b[2] + c;
                    if (!inb(i, j, k)) continue;
                    V3<T> xn = xnode(i, j, k);
                    V3<T> dgp{{xn[0] - x[0], xn[1] - x[1], xn[2] - x[2]}};
                    T u0 = dgp[0] / dx, u1 = dgp[1] / dx, u2 = dgp[2] / dx;
                    T n0 = bs(p->order, u0), n1 = bs(p->order, u1), n2 = bs(p->order, u2);
                    T w = n0 * n1 * n2;
                    V3<T> gw{{0, 0, 0}};
                    if (grad) {
                        T d0 = dbs(p->order, u0), d1 = dbs(p->order, u1), d2 = dbs(p->order, u2);
                        T s = T(-1) / dx;
                        gw = V3<T>{{s * (d0 * n1 * n2), s * (n0 * d1 * n2), s * (n0 * n1 * d2)}};
                    }
                    fn(id(i, j, k), i, j, k, dgp, w, gw);
                }
    }
};

template <typename T>
struct State {
    int n;
    const T *x, *v, *F, *C, *dFc, *m, *vol, *lam, *mu;
    V3<T> X(int p) const { return ldv(x + 3 * p); }
    V3<T> Vv(int p) const { return ldv(v + 3 * p); }
    M3<T> Fm(int p) const { return ld(F + 9 * p); }
    M3<T> Cm(int p) const { return ld(C + 9 * p); }
    M3<T> Ft(int p) const { return dFc ? add(Fm(p), ld(dFc + 9 * p)) : Fm(p); }
};

// Grid after P2G + grid update of one step.
template <typename T>
struct GridState {
    std::vector<T> m, p, v;  // m (G), p (G,3), v (G,3)
};

// P2G (ForwardSimulation.cpp:26-80, 201-207) with stress P from P_op_1
// (:16-24, 188-194), then the mass clamp (:165-170) and the grid velocity
// update (:82-87, 209-215), with sticky walls when enabled.
template <typename T>
void p2g_grid(const Ctx<T>& cx, const State<T>& s, const std::vector<M3<T>>& P, GridState<T>& g) {
    const orc_params* prm = cx.p;
    const T dt = (T)prm->dt, drag = (T)prm->drag, zeta = (T)prm->zeta;
    size_t G = cx.G();
    g.m.assign(G, T(0));
    g.p.assign(3 * G, T(0));
    g.v.assign(3 * G, T(0));
    for (int q = 0; q < s.n; ++q) {
        V3<T> x = s.X(q), v = s.Vv(q);
        M3<T> Ft = s.Ft(q), C = s.Cm(q);
        T m = s.m[q], vol = s.vol[q];
        T tau = zeta * trace(C);
        M3<T> stress = mul(P[q], tr(Ft));
        stress(0,0) += tau; stress(1,1) += tau; stress(2,2) += tau;
        M3<T> A = add(scl(-cx.dinv * dt * vol, stress), scl(m, C));
        V3<T> mv{{m * v[0] * (T(1) - dt * drag), m * v[1] * (T(1) - dt * drag), m * v[2] * (T(1) - dt * drag)}};
        cx.for_stencil(x, false, [&](size_t id, int, int, int, const V3<T>& dgp, T w, const V3<T>&) {
            g.m[id] += w * m;
            V3<T> Ad = mulv(A, dgp);
            for (int a = 0; a < 3; ++a) g.p[3 * id + a] += w * (mv[a] + Ad[a]);
        });
    }
    for (int i = 0; i < cx.nx; ++i)
        for (int j = 0; j < cx.ny; ++j)
            for (int k = 0; k < cx.nz; ++k) {
                size_t id = cx.id(i, j, k);
                if (g.m[id] < T(0)) g.m[id] = T(0);
                if (g.m[id] > T(1e-12) && !cx.wall(i, j, k))
                    for (int a = 0; a < 3; ++a)
                        g.v[3 * id + a] = g.p[3 * id + a] / g.m[id] + dt * (T)prm->f_ext[a];
            }
}

// Per-particle forward results needed again by the adjoint.
template <typename T>
struct PartFwd {
    V3<T> vn;      // v'
    M3<T> Craw;    // C' before the trust region
    M3<T> Cn;      // C' after the trust region
    T sJ;          // trust-region scale applied (1 when inactive)
    int tr_state;  // 0 inactive, 1 scaled (derivative live), 2 clamped to 0
    T denom;       // min(-1e-6, dt trC) when active
    T J0;
    V3<T> xn;
    M3<T> Fn;
};

// G2P (ForwardSimulation.cpp:89-123) and P_op_2 with the trust region and
// F smoothing (:131-157).
template <typename T>
PartFwd<T> g2p_update(const Ctx<T>& cx, const State<T>& s, const GridState<T>& g, int q) {
    const orc_params* prm = cx.p;
    const T dt = (T)prm->dt;
    PartFwd<T> r;
    V3<T> x = s.X(q);
    r.vn = V3<T>{{0, 0, 0}};
    r.Craw = zero3<T>();
    cx.for_stencil(x, false, [&](size_t id, int, int, int, const V3<T>& dgp, T w, const V3<T>&) {
        V3<T> vi{{g.v[3 * id], g.v[3 * id + 1], g.v[3 * id + 2]}};
        for (int a = 0; a < 3; ++a) r.vn[a] += w * vi[a];
        T cw = cx.dinv * w;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) r.Craw(a, b) += cw * vi[a] * dgp[b];
    });
    M3<T> Ft = s.Ft(q);
    r.J0 = std::max(det3(Ft), T(1e-6));
    r.Cn = r.Craw;
    r.sJ = T(1);
    r.tr_state = 0;
    r.denom = T(0);
    if (prm->trust_region) {
        T trC = trace(r.Craw);
        if (trC < T(0)) {
            T denom = std::min(T(-1e-6), dt * trC);
            T sv = ((T)prm->jmin / r.J0 - T(1)) / denom;
            if (sv < T(1)) {
                r.tr_state = sv < T(0) ? 2 : 1;
                sv = std::clamp(sv, T(0), T(1));
                r.sJ = sv;
                r.denom = denom;
                r.Cn = scl(sv, r.Craw);
            }
        }
    }
    T sm = (T)prm->smoothing;
    M3<T> Fn = mul(add(eye3<T>(), scl(dt, r.Cn)), Ft);
    r.Fn = add(scl(T(1) - sm, Fn), scl(sm, s.Fm(q)));
    r.xn = V3<T>{{x[0] + dt * r.vn[0], x[1] + dt * r.vn[1], x[2] + dt * r.vn[2]}};
    return r;
}

template <typename T>
void stress_all(const orc_params* prm, const State<T>& s, std::vector<M3<T>>& P) {
    P.resize(s.n);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < s.n; ++q) P[q] = pk1(prm->model, s.Ft(q), s.lam[q], s.mu[q]);
}

template <typename T>
void forward_step(const orc_params* prm, const State<T>& s, T* ox, T* ov, T* oF, T* oC, T* oP, T* gm, T* gv) {
    Ctx<T> cx(prm);
    std::vector<M3<T>> P;
    stress_all(prm, s, P);
    GridState<T> g;
    p2g_grid(cx, s, P, g);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < s.n; ++q) {
        PartFwd<T> r = g2p_update(cx, s, g, q);
        stv(r.xn, ox + 3 * q);
        stv(r.vn, ov + 3 * q);
        st(r.Fn, oF + 9 * q);
        st(r.Cn, oC + 9 * q);
        if (oP) st(P[q], oP + 9 * q);
    }
    if (gm) std::copy(g.m.begin(), g.m.end(), gm);
    if (gv) std::copy(g.v.begin(), g.v.end(), gv);
}

// ---------------------------------------------------------------------------
// Adjoint of one step.  adjoint == 1 restates Back_Timestep
// (BackPropagation.cpp:9-144) formula by formula; adjoint == 0 is the exact
// vector-Jacobian product of forward_step above.
template <typename T>
void backward_step(const orc_params* prm, const State<T>& s, const T* gx1, const T* gv1, const T* gF1, const T* gC1,
                   T* gx, T* gv, T* gF, T* gC, T* g_dFc, T* gP_out) {
    Ctx<T> cx(prm);
    const T dt = (T)prm->dt, drag = (T)prm->drag, zeta = (T)prm->zeta, sm = (T)prm->smoothing;
    const int n = s.n;
    const bool ref = prm->adjoint == 1;
    std::vector<M3<T>> P;
    stress_all(prm, s, P);
    GridState<T> g;
    p2g_grid(cx, s, P, g);
    std::vector<PartFwd<T>> fw(n);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) fw[q] = g2p_update(cx, s, g, q);

    // Phase 1 (BackPropagation.cpp:22-32).
    std::vector<V3<T>> gvtot(n);
    std::vector<M3<T>> gCraw(n), gFt(n), gFdir(n);
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) {
        M3<T> G1 = ld(gF1 + 9 * q), gCn = ld(gC1 + 9 * q);
        V3<T> X1 = ldv(gx1 + 3 * q), V1 = ldv(gv1 + 3 * q);
        M3<T> Ft = s.Ft(q);
        M3<T> A1 = add(eye3<T>(), scl(dt, fw[q].Cn));
        for (int a = 0; a < 3; ++a) gvtot[q][a] = dt * X1[a] + V1[a];
        if (ref) {
            M3<T> dF = mul(tr(A1), G1);
            gFt[q] = add(scl(T(1) - sm, dF), scl(sm, G1));  // first (1-s)A^T G + s G
            gFdir[q] = zero3<T>();
            gCraw[q] = add(scl(dt, mul(G1, tr(Ft))), gCn);
        } else {
            gFt[q] = scl(T(1) - sm, mul(tr(A1), G1));
            gFdir[q] = scl(sm, G1);
            M3<T> gCfin = add(gCn, scl((T(1) - sm) * dt, mul(G1, tr(Ft))));
            if (fw[q].tr_state == 0) {
                gCraw[q] = gCfin;
            } else if (fw[q].tr_state == 2) {
                gCraw[q] = zero3<T>();
            } else {
                T sJ = fw[q].sJ;
                T cc = ddot(gCfin, fw[q].Craw);
                M3<T> gc = scl(sJ, gCfin);
                T trC = trace(fw[q].Craw);
                if (dt * trC < T(-1e-6)) {
                    T dsdtr = -sJ / trC;
                    gc(0,0) += cc * dsdtr; gc(1,1) += cc * dsdtr; gc(2,2) += cc * dsdtr;
                }
                gCraw[q] = gc;
                // s depends on J0 = max(det Ft, 1e-6)
                T detFt = det3(Ft);
                if (detFt > T(1e-6)) {
                    T dsdJ = -(T)prm->jmin / (fw[q].J0 * fw[q].J0) / fw[q].denom;
                    gFt[q] = add(gFt[q], scl(cc * dsdJ, cof3(Ft)));
                }
            }
        }
    }

    // Phase 2: G2P adjoint scatter (BackPropagation.cpp:36-55), serial.
    size_t G = cx.G();
    std::vector<T> gvg(3 * G, T(0));
    for (int q = 0; q < n; ++q) {
        V3<T> x = s.X(q);
        cx.for_stencil(x, false, [&](size_t id, int, int, int, const V3<T>& dgp, T w, const V3<T>&) {
            V3<T> cd = mulv(gCraw[q], dgp);
            for (int a = 0; a < 3; ++a) gvg[3 * id + a] += gvtot[q][a] * w + cx.dinv * w * cd[a];
        });
    }
    // Phase 3: grid adjoint (BackPropagation.cpp:59-72).
    std::vector<T> gpg(3 * G, T(0)), gmg(G, T(0));
    for (int i = 0; i < cx.nx; ++i)
        for (int j = 0; j < cx.ny; ++j)
            for (int k = 0; k < cx.nz; ++k) {
                size_t id = cx.id(i, j, k);
                T mi = g.m[id];
                if (!(mi > T(1e-12))) continue;
                if (cx.wall(i, j, k)) continue;  // wall nodes: v = 0, no adjoint
                V3<T> vi{{g.v[3 * id], g.v[3 * id + 1], g.v[3 * id + 2]}};
                if (!ref) for (int a = 0; a < 3; ++a) vi[a] -= dt * (T)prm->f_ext[a];
                T vd = vi[0] * gvg[3 * id] + vi[1] * gvg[3 * id + 1] + vi[2] * gvg[3 * id + 2];
                for (int a = 0; a < 3; ++a) gpg[3 * id + a] = gvg[3 * id + a] / mi;
                gmg[id] = T(-1) / mi * vd;
            }
    // Phase 4: P2G adjoint gather + stress VJP (BackPropagation.cpp:76-143).
#pragma omp parallel for schedule(static)
    for (int q = 0; q < n; ++q) {
        V3<T> x = s.X(q), v = s.Vv(q);
        M3<T> Ft = s.Ft(q), C = s.Cm(q), Ftt = tr(Ft);
        T m = s.m[q], vol = s.vol[q];
        const T k0 = cx.dinv * dt * vol;
        M3<T> Gm = add(scl(-k0, mul(P[q], Ftt)), scl(m, C));
        T bulk = -k0 * (zeta * trace(C));
        M3<T> Gt = Gm;
        Gt(0,0) += bulk; Gt(1,1) += bulk; Gt(2,2) += bulk;
        V3<T> dx{{0, 0, 0}}, dv{{0, 0, 0}};
        M3<T> dP = zero3<T>(), dFa = zero3<T>(), dC = zero3<T>(), gA = zero3<T>();
        const T mom = ref ? m : m * (T(1) - dt * drag);
        cx.for_stencil(x, true, [&](size_t id, int, int, int, const V3<T>& dgp, T w, const V3<T>& gw) {
            V3<T> h{{gpg[3 * id], gpg[3 * id + 1], gpg[3 * id + 2]}};
            V3<T> vi{{g.v[3 * id], g.v[3 * id + 1], g.v[3 * id + 2]}};
            T gm = gmg[id];
            if (ref) {
                // :112-114, :117-123 (bulk on the diagonal only), :126-136
                dP = sub(dP, scl(w * cx.dinv * dt * vol, outer(h, mulv(Ftt, dgp))));
                dFa = sub(dFa, scl(w * cx.dinv * dt * vol, outer(dgp, mulv(tr(P[q]), h))));
                dC = add(dC, scl(w * m, outer(h, dgp)));
                M3<T> o = scl(w, outer(h, dgp));
                const T factor = -cx.dinv * dt * vol * zeta;
                dC(0,0) += factor * o(0,0); dC(1,1) += factor * o(1,1); dC(2,2) += factor * o(2,2);
            } else {
                gA = add(gA, scl(w, outer(h, dgp)));
            }
            // dL/dx from mass, momentum, v' and C' (:126-134)
            V3<T> Gd = mulv(Gt, dgp);
            V3<T> qv{{mom * v[0] + Gd[0], mom * v[1] + Gd[1], mom * v[2] + Gd[2]}};
            T qh = dotv(qv, h);
            V3<T> Gth = mulv(tr(Gt), h);
            T vg = dotv(vi, gvtot[q]);
            T cc = ddot(gCraw[q], outer(vi, dgp));
            V3<T> Ctv = mulv(tr(gCraw[q]), vi);
            for (int a = 0; a < 3; ++a) {
                dx[a] += m * gm * gw[a];
                dx[a] += gw[a] * qh - w * Gth[a];
                dx[a] += gw[a] * vg;
                dx[a] += cx.dinv * (cc * gw[a] - w * Ctv[a]);
            }
            for (int a = 0; a < 3; ++a) dv[a] += w * m * (T(1) - dt * drag) * h[a];
        });
        M3<T> gFt_q = gFt[q];
        if (ref) {
            M3<T> A1 = add(eye3<T>(), scl(dt, fw[q].Cn));
            gFt_q = add(gFt_q, dFa);
            gFt_q = add(gFt_q, mul(tr(A1), ld(gF1 + 9 * q)));  // :140, added again
            gFt_q = add(gFt_q, pk1_hvp(prm->model, Ft, s.lam[q], s.mu[q], dP));
        } else {
            dP = scl(-k0, mul(gA, Ft));
            gFt_q = add(gFt_q, scl(-k0, mul(tr(gA), P[q])));
            dC = scl(m, gA);
            T tb = -k0 * zeta * trace(gA);
            dC(0,0) += tb; dC(1,1) += tb; dC(2,2) += tb;
            gFt_q = add(gFt_q, pk1_hvp(prm->model, Ft, s.lam[q], s.mu[q], dP));
        }
        V3<T> X1 = ldv(gx1 + 3 * q);
        for (int a = 0; a < 3; ++a) dx[a] += X1[a];
        stv(dx, gx + 3 * q);
        stv(dv, gv + 3 * q);
        st(add(gFt_q, gFdir[q]), gF + 9 * q);
        st(dC, gC + 9 * q);
        if (g_dFc) st(ref ? add(gFt_q, gFdir[q]) : gFt_q, g_dFc + 9 * q);
        if (gP_out) st(dP, gP_out + 9 * q);
    }
}

// EndLayerMassLoss (CompGraph.cpp:42-129).
template <typename T>
double mass_loss(const orc_params* prm, int n, const T* x, const T* m, const T* target_m, T* gx, T* gm_out, T* gdm_out) {
    Ctx<T> cx(prm);
    size_t G = cx.G();
    std::vector<T> gm(G, T(0));
    for (int q = 0; q < n; ++q) {
        V3<T> xp = ldv(x + 3 * q);
        cx.for_stencil(xp, false, [&](size_t id, int, int, int, const V3<T>&, T w, const V3<T>&) { gm[id] += w * m[q]; });
    }
    for (auto& e : gm) if (e < T(0)) e = T(0);
    T sum_c = 0, sum_t = 0;
    for (size_t i = 0; i < G; ++i) { sum_c += gm[i]; sum_t += target_m[i]; }
    const T sc = (sum_t + T(1e-6)) / (sum_c + T(1e-6));
    const T eps = T(1e-4), min_mass = T(1e-3);
    double loss = 0;
    std::vector<T> dLdm(G);
    for (size_t i = 0; i < G; ++i) {
        T cms = sc * gm[i];
        T ld_ = std::log(cms + T(1) + eps) - std::log(target_m[i] + T(1) + eps);
        loss += (double)(T(0.5) * ld_ * ld_);
        dLdm[i] = (sc / (cms + T(1) + eps)) * ld_;
        if (cms < min_mass) {
            T diff = min_mass - cms;
            loss += (double)(diff * diff);
            dLdm[i] += T(-2) * diff * sc;
        }
    }
    for (int q = 0; q < n; ++q) {
        V3<T> xp = ldv(x + 3 * q);
        V3<T> acc{{0, 0, 0}};
        cx.for_stencil(xp, true, [&](size_t id, int, int, int, const V3<T>&, T, const V3<T>& gw) {
            T pen = target_m[id] > T(1e-12) ? T(1) : T(5);
            for (int a = 0; a < 3; ++a) acc[a] += pen * m[q] * dLdm[id] * gw[a];
        });
        stv(acc, gx + 3 * q);
    }
    if (gm_out) std::copy(gm.begin(), gm.end(), gm_out);
    if (gdm_out) std::copy(dLdm.begin(), dLdm.end(), gdm_out);
    return loss;
}

// P2G masses with dt = 0 (bind.cpp:546-549).
template <typename T>
void p2g_mass(const orc_params* prm, int n, const T* x, const T* m, T* grid_m) {
    Ctx<T> cx(prm);
    std::fill(grid_m, grid_m + cx.G(), T(0));
    for (int q = 0; q < n; ++q) {
        V3<T> xp = ldv(x + 3 * q);
        cx.for_stencil(xp, false, [&](size_t id, int, int, int, const V3<T>&, T w, const V3<T>&) { grid_m[id] += w * m[q]; });
    }
}

// CalculatePointCloudVolumes (ForwardSimulation.cpp:234-271).
template <typename T>
void volumes(const orc_params* prm, int n, const T* x, const T* m, T* vol, T* grid_m) {
    Ctx<T> cx(prm);
    std::vector<T> gm(cx.G());
    p2g_mass(prm, n, x, m, gm.data());
    const T cell = cx.dx * cx.dx * cx.dx;
    for (int q = 0; q < n; ++q) {
        V3<T> xp = ldv(x + 3 * q);
        T mg = 0;
        cx.for_stencil(xp, false, [&](size_t id, int, int, int, const V3<T>&, T w, const V3<T>&) { mg += w * gm[id]; });
        T rho = mg / std::max(cell, T(1e-12));
        vol[q] = rho > std::numeric_limits<T>::epsilon() ? m[q] / rho : T(0);
    }
    if (grid_m) std::copy(gm.begin(), gm.end(), grid_m);
}

// Descend_Adam (PointCloud.cpp:83-147): AMSGrad moments on the normalised
// gradient, absolute/relative step cap, and a damped update near stationarity.
template <typename T>
void adam(int n, const T* dLdF, T* dFc, T* mom, T* vec, T* vmax, double alpha_, double gn, double b1_, double b2_, double eps_, int timestep) {
    const T alpha = (T)alpha_, b1 = (T)b1_, b2 = (T)b2_, eps = (T)eps_;
    const T inv_gn = T(1) / std::max((T)gn, T(1e-12));
    const int t = std::max(1, timestep);
    const T c1 = T(1) - (T)std::pow((double)b1, t);
    const T c2 = T(1) - (T)std::pow((double)b2, t);
    for (int q = 0; q < n; ++q) {
        M3<T> g = scl(inv_gn, ld(dLdF + 9 * q));
        M3<T> M = ld(mom + 9 * q), Vv = ld(vec + 9 * q), Vm = ld(vmax + 9 * q), X = ld(dFc + 9 * q);
        M3<T> step;
        for (int i = 0; i < 9; ++i) {
            M.a[i] = b1 * M.a[i] + (T(1) - b1) * g.a[i];
            Vv.a[i] = b2 * Vv.a[i] + (T(1) - b2) * g.a[i] * g.a[i];
            T mh = M.a[i] / c1, vh = Vv.a[i] / c2;
            Vm.a[i] = std::max(Vm.a[i], vh);
            T ds = std::max(std::sqrt(Vm.a[i]), T(1e-8));
            step.a[i] = mh / (ds + eps);
        }
        T step_norm = fnorm(step);
        T x_norm = fnorm(X);
        T cap = std::max(T(5e-3), T(5e-2) * std::max(T(1), x_norm));
        M3<T> st_ = step;
        if (step_norm > T(0)) {
            T scaled = alpha * step_norm;
            if (scaled > cap) st_ = scl(cap / scaled, st_);
        }
        M3<T> Xold = X;
        X = sub(X, scl(alpha, st_));
        T small_cap = std::max(T(1e-4), T(1e-3) * std::max(T(1), fnorm(Xold)));
        if (alpha * step_norm <= small_cap) X = add(Xold, scl(T(0.2), sub(X, Xold)));
        st(X, dFc + 9 * q);
        st(M, mom + 9 * q);
        st(Vv, vec + 9 * q);
        st(Vm, vmax + 9 * q);
    }
}

template <typename T>
State<T> mk(int n, const T* x, const T* v, const T* F, const T* C, const T* dFc, const T* m, const T* vol, const T* lam, const T* mu) {
    return State<T>{n, x, v, F, C, dFc, m, vol, lam, mu};
}

}  // namespace

#define ORC_DEF(S, T)                                                                                      \
    void orc_forward_step_##S(const orc_params* prm, int n, const T* x, const T* v, const T* F,            \
                              const T* C, const T* dFc, const T* m, const T* vol, const T* lam,           \
                              const T* mu, T* ox, T* ov, T* oF, T* oC, T* oP, T* gm, T* gv) {             \
        forward_step<T>(prm, mk<T>(n, x, v, F, C, dFc, m, vol, lam, mu), ox, ov, oF, oC, oP, gm, gv);    \
    }                                                                                                      \
    void orc_backward_step_##S(const orc_params* prm, int n, const T* x, const T* v, const T* F,           \
                               const T* C, const T* dFc, const T* m, const T* vol, const T* lam,          \
                               const T* mu, const T* gx1, const T* gv1, const T* gF1, const T* gC1,       \
                               T* gx, T* gv, T* gF, T* gC, T* g_dFc, T* gP) {                              \
        backward_step<T>(prm, mk<T>(n, x, v, F, C, dFc, m, vol, lam, mu), gx1, gv1, gF1, gC1, gx, gv,    \
                         gF, gC, g_dFc, gP);                                                               \
    }                                                                                                      \
    void orc_pk1_##S(int model, int n, const T* F, const T* lam, const T* mu, T* P) {                      \
        for (int q = 0; q < n; ++q) st(pk1<T>(model, ld(F + 9 * q), lam[q], mu[q]), P + 9 * q);           \
    }                                                                                                      \
    void orc_pk1_hvp_##S(int model, int n, const T* F, const T* lam, const T* mu, const T* dF, T* out) {  \
        for (int q = 0; q < n; ++q)                                                                        \
            st(pk1_hvp<T>(model, ld(F + 9 * q), lam[q], mu[q], ld(dF + 9 * q)), out + 9 * q);              \
    }                                                                                                      \
    void orc_energy_##S(int model, int n, const T* F, const T* lam, const T* mu, T* psi) {                 \
        for (int q = 0; q < n; ++q) psi[q] = energy<T>(model, ld(F + 9 * q), lam[q], mu[q]);               \
    }                                                                                                      \
    void orc_svd_##S(int n, const T* F, T* U, T* sigma, T* V) {                                            \
        for (int q = 0; q < n; ++q) {                                                                      \
            M3<double> u, vv;                                                                              \
            double s[3];                                                                                   \
            svd3(cast<T, double>(ld(F + 9 * q)), u, s, vv);                                                \
            st(cast<double, T>(u), U + 9 * q);                                                             \
            st(cast<double, T>(vv), V + 9 * q);                                                            \
            for (int k = 0; k < 3; ++k) sigma[3 * q + k] = (T)s[k];                                        \
        }                                                                                                  \
    }                                                                                                      \
    void orc_polar_##S(int n, const T* F, T* R, T* Sm) {                                                   \
        for (int q = 0; q < n; ++q) {                                                                      \
            M3<T> r, s;                                                                                    \
            polar3<T>(ld(F + 9 * q), r, s);                                                                \
            st(r, R + 9 * q);                                                                              \
            st(s, Sm + 9 * q);                                                                             \
        }                                                                                                  \
    }                                                                                                      \
    void orc_p2g_mass_##S(const orc_params* prm, int n, const T* x, const T* m, T* grid_m) {              \
        p2g_mass<T>(prm, n, x, m, grid_m);                                                                 \
    }                                                                                                      \
    void orc_volumes_##S(const orc_params* prm, int n, const T* x, const T* m, T* vol, T* grid_m) {        \
        volumes<T>(prm, n, x, m, vol, grid_m);                                                             \
    }                                                                                                      \
    double orc_mass_loss_##S(const orc_params* prm, int n, const T* x, const T* m, const T* target_m,      \
                             T* gx, T* grid_m, T* grid_dLdm) {                                             \
        return mass_loss<T>(prm, n, x, m, target_m, gx, grid_m, grid_dLdm);                               \
    }                                                                                                      \
    void orc_adam_##S(int n, const T* dLdF, T* dFc, T* mom, T* vec, T* vmax, double alpha,                \
                      double grad_norm, double beta1, double beta2, double eps, int timestep) {            \
        adam<T>(n, dLdF, dFc, mom, vec, vmax, alpha, grad_norm, beta1, beta2, eps, timestep);              \
    }                                                                                                      \
    double orc_dLdF_norm_##S(int n, const T* dLdF) {                                                       \
        double s = 0;                                                                                      \
        for (int i = 0; i < 9 * n; ++i) s += (double)dLdF[i] * (double)dLdF[i];                           \
        return std::sqrt(s);                                                                               \
    }                                                                                                      \
    void orc_bspline_##S(int order, int n, const T* u, T* w, T* dw) {                                      \
        for (int i = 0; i < n; ++i) {                                                                      \
            w[i] = bs<T>(order, u[i]);                                                                     \
            dw[i] = dbs<T>(order, u[i]);                                                                   \
        }                                                                                                  \
    }

extern "C" {
ORC_DEF(f32, float)
ORC_DEF(f64, double)
}
