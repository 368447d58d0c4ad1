// Device-side 3x3 algebra, B-spline weights, 3x3 SVD/polar decomposition and
// the two constitutive models with their Hessian actions.
//
// Everything is templated on the storage/arithmetic type T (float or double).
// The determinant, its log and the inverse-transpose used by the stress and
// its derivative are always evaluated in double ("W" below) so that the FP32
// path keeps FP64 accuracy exactly where compression makes them sensitive.
#pragma once
#include <hip/hip_runtime.h>

#define DMPM_DI __device__ __forceinline__

namespace dmpm {

template <typename T> struct m3 { T a[9]; };  // row-major: (r,c) at a[3r+c]
template <typename T> struct v3 { T a[3]; };

template <typename T> DMPM_DI m3<T> mzero() { m3<T> M;
#pragma unroll
    for (int i = 0; i < 9; ++i) M.a[i] = T(0); return M; }
template <typename T> DMPM_DI m3<T> meye() { m3<T> M = mzero<T>(); M.a[0] = M.a[4] = M.a[8] = T(1); return M; }
template <typename T> DMPM_DI v3<T> vzero() { v3<T> v; v.a[0] = v.a[1] = v.a[2] = T(0); return v; }

template <typename T> DMPM_DI m3<T> madd(const m3<T>& A, const m3<T>& B) { m3<T> R;
#pragma unroll
    for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] + B.a[i]; return R; }
template <typename T> DMPM_DI m3<T> msub(const m3<T>& A, const m3<T>& B) { m3<T> R;
#pragma unroll
    for (int i = 0; i < 9; ++i) R.a[i] = A.a[i] - B.a[i]; return R; }
template <typename T> DMPM_DI m3<T> mscale(T s, const m3<T>& A) { m3<T> R;
#pragma unroll
    for (int i = 0; i < 9; ++i) R.a[i] = s * A.a[i]; return R; }
template <typename T> DMPM_DI m3<T> mtr(const m3<T>& A) { m3<T> R;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R.a[3 * r + c] = A.a[3 * c + r];
    return R; }
template <typename T> DMPM_DI m3<T> mmul(const m3<T>& A, const m3<T>& B) { m3<T> R;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            R.a[3 * r + c] = A.a[3 * r] * B.a[c] + A.a[3 * r + 1] * B.a[3 + c] + A.a[3 * r + 2] * B.a[6 + c];
    return R; }
// A^T B
template <typename T> DMPM_DI m3<T> mtmul(const m3<T>& A, const m3<T>& B) { m3<T> R;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            R.a[3 * r + c] = A.a[r] * B.a[c] + A.a[3 + r] * B.a[3 + c] + A.a[6 + r] * B.a[6 + c];
    return R; }
// A B^T
template <typename T> DMPM_DI m3<T> mmult(const m3<T>& A, const m3<T>& B) { m3<T> R;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            R.a[3 * r + c] = A.a[3 * r] * B.a[3 * c] + A.a[3 * r + 1] * B.a[3 * c + 1] + A.a[3 * r + 2] * B.a[3 * c + 2];
    return R; }
template <typename T> DMPM_DI v3<T> mv(const m3<T>& A, const v3<T>& x) { v3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) r.a[i] = A.a[3 * i] * x.a[0] + A.a[3 * i + 1] * x.a[1] + A.a[3 * i + 2] * x.a[2];
    return r; }
// A^T x
template <typename T> DMPM_DI v3<T> mtv(const m3<T>& A, const v3<T>& x) { v3<T> r;
#pragma unroll
    for (int i = 0; i < 3; ++i) r.a[i] = A.a[i] * x.a[0] + A.a[3 + i] * x.a[1] + A.a[6 + i] * x.a[2];
    return r; }
template <typename T> DMPM_DI T mddot(const m3<T>& A, const m3<T>& B) { T s = T(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) s += A.a[i] * B.a[i]; return s; }
template <typename T> DMPM_DI T vdot(const v3<T>& a, const v3<T>& b) { return a.a[0] * b.a[0] + a.a[1] * b.a[1] + a.a[2] * b.a[2]; }
template <typename T> DMPM_DI T mtrace(const m3<T>& A) { return A.a[0] + A.a[4] + A.a[8]; }
template <typename T> DMPM_DI m3<T> vouter(const v3<T>& a, const v3<T>& b) { m3<T> R;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) R.a[3 * r + c] = a.a[r] * b.a[c];
    return R; }
template <typename U, typename T> DMPM_DI m3<U> mcast(const m3<T>& A) { m3<U> R;
#pragma unroll
    for (int i = 0; i < 9; ++i) R.a[i] = (U)A.a[i]; return R; }

template <typename T> DMPM_DI T mdet(const m3<T>& m) {
    return m.a[0] * (m.a[4] * m.a[8] - m.a[5] * m.a[7])
         - m.a[1] * (m.a[3] * m.a[8] - m.a[5] * m.a[6])
         + m.a[2] * (m.a[3] * m.a[7] - m.a[4] * m.a[6]);
}
// cofactor matrix cof(A) = det(A) A^{-T}
template <typename T> DMPM_DI m3<T> mcof(const m3<T>& m) {
    m3<T> c;
    c.a[0] = m.a[4] * m.a[8] - m.a[5] * m.a[7];
    c.a[1] = m.a[5] * m.a[6] - m.a[3] * m.a[8];
    c.a[2] = m.a[3] * m.a[7] - m.a[4] * m.a[6];
    c.a[3] = m.a[2] * m.a[7] - m.a[1] * m.a[8];
    c.a[4] = m.a[0] * m.a[8] - m.a[2] * m.a[6];
    c.a[5] = m.a[1] * m.a[6] - m.a[0] * m.a[7];
    c.a[6] = m.a[1] * m.a[5] - m.a[2] * m.a[4];
    c.a[7] = m.a[2] * m.a[3] - m.a[0] * m.a[5];
    c.a[8] = m.a[0] * m.a[4] - m.a[1] * m.a[3];
    return c;
}

// ---------------------------------------------------------------------------
// B-spline weights along one axis.  fx = (x - min)/dx in cell units.
// Cubic (reference, Interpolation.h:5-30): base = floor(fx) - 1, 4 nodes,
//   N(u) = |u|^3/2 - u^2 + 2/3 (|u|<1), (2-|u|)^3/6 (1<=|u|<2).
// Quadratic: base = floor(fx - 1/2), 3 nodes,
//   N(u) = 3/4 - u^2 (|u|<1/2), (3/2-|u|)^2/2 (1/2<=|u|<3/2).
// u_k = (base + k) - fx is (x_node - x_p)/dx; w[k] = N(u_k), dw[k] = N'(u_k).
// Each k lies in one fixed piece of N, so the pieces are written out per k.
template <int ORDER> struct Spline;

template <> struct Spline<2> {
    static constexpr int K = 3;
    template <typename T>
    static DMPM_DI int weights(T fx, T* u, T* w, T* dw) {
        int b = (int)floor(fx - T(0.5));
        T d = fx - (T)b;  // in [0.5, 1.5)
        T e0 = T(1.5) - d, e1 = T(1) - d, e2 = d - T(0.5);
        u[0] = -d; u[1] = e1; u[2] = T(2) - d;
        w[0] = T(0.5) * e0 * e0;
        w[1] = T(0.75) - e1 * e1;
        w[2] = T(0.5) * e2 * e2;
        dw[0] = e0;
        dw[1] = T(-2) * e1;
        dw[2] = -e2;
        return b;
    }
    template <typename T> static DMPM_DI T dinv(T dx) { return T(4) / (dx * dx); }
};

template <> struct Spline<3> {
    static constexpr int K = 4;
    template <typename T>
    static DMPM_DI int weights(T fx, T* u, T* w, T* dw) {
        int b = (int)floor(fx) - 1;
        T d = fx - (T)b;  // in [1, 2)
        T a = d - T(1);   // |u_1|, in [0,1)
        T c = T(2) - d;   // |u_2| = u_2, in (0,1]
        T e = T(2) - d;   // 2 - |u_0|
        u[0] = -d; u[1] = -a; u[2] = c; u[3] = T(3) - d;
        w[0] = e * e * e / T(6);
        w[1] = T(0.5) * a * a * a - a * a + T(2) / T(3);
        w[2] = T(0.5) * c * c * c - c * c + T(2) / T(3);
        w[3] = a * a * a / T(6);
        dw[0] = T(0.5) * e * e;
        dw[1] = T(-1.5) * a * a + T(2) * a;
        dw[2] = T(1.5) * c * c - T(2) * c;
        dw[3] = T(-0.5) * a * a;
        return b;
    }
    template <typename T> static DMPM_DI T dinv(T dx) { return T(3) / (dx * dx); }
};

// ---------------------------------------------------------------------------
// 3x3 SVD F = U diag(s) V^T with U, V rotations, s sorted by decreasing
// magnitude and only s[2] allowed to be negative (the convention of the
// reference's polar decomposition, ImplicitQRSVD.h:852-875).
// Method: cyclic Jacobi eigen-decomposition of F^T F (fixed sweep count, no
// data-dependent loop), sort, then U from a Gram-Schmidt QR of F V; the sign
// of s[2] comes out of the QR.  Following the structure of McAdams et al.,
// "Computing the SVD of 3x3 matrices with minimal branching and elementary
// floating point operations" (2011), with exact Jacobi angles.
template <typename T>
DMPM_DI void jacobi_rot(T* S, m3<T>& V, int p, int q) {
    // S symmetric stored full row-major
    T spq = S[3 * p + q];
    T spp = S[3 * p + p], sqq = S[3 * q + q];
    T tiny = T(1e-30);
    if (fabs(spq) <= tiny * (fabs(spp) + fabs(sqq)) + T(1e-36)) return;
    T theta = (sqq - spp) / (T(2) * spq);
    T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + sqrt(theta * theta + T(1)));
    T c = T(1) / sqrt(t * t + T(1));
    T s = t * c;
    // S <- J^T S J, J = rotation in (p,q): columns p,q mixed
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T skp = S[3 * k + p], skq = S[3 * k + q];
        S[3 * k + p] = c * skp - s * skq;
        S[3 * k + q] = s * skp + c * skq;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T spk = S[3 * p + k], sqk = S[3 * q + k];
        S[3 * p + k] = c * spk - s * sqk;
        S[3 * q + k] = s * spk + c * sqk;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T vkp = V.a[3 * k + p], vkq = V.a[3 * k + q];
        V.a[3 * k + p] = c * vkp - s * vkq;
        V.a[3 * k + q] = s * vkp + c * vkq;
    }
}

template <typename T>
DMPM_DI void swapcols(m3<T>& V, T* e, int i, int j) {
    T t = e[i]; e[i] = e[j]; e[j] = t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T a = V.a[3 * k + i];
        V.a[3 * k + i] = V.a[3 * k + j];
        V.a[3 * k + j] = -a;  // keep det V = +1
    }
}

template <typename T>
DMPM_DI void svd3(const m3<T>& F, m3<T>& U, v3<T>& sig, m3<T>& V) {
    m3<T> S = mtmul(F, F);
    V = meye<T>();
#pragma unroll 1
    for (int sweep = 0; sweep < 6; ++sweep) {
        jacobi_rot(S.a, V, 0, 1);
        jacobi_rot(S.a, V, 0, 2);
        jacobi_rot(S.a, V, 1, 2);
    }
    // Sort by adjacent swaps only, so that equal eigenvalues keep their
    // order: with |sigma_2| = |sigma_3| and det F < 0 the closest rotation is
    // not unique, and this order gives the reference's choice
    // (diag(0.5, -0.5, 1) -> R = I, as ImplicitQRSVD.h's polar factor).
    T e[3] = {S.a[0], S.a[4], S.a[8]};
    if (e[0] < e[1]) swapcols(V, e, 0, 1);
    if (e[1] < e[2]) swapcols(V, e, 1, 2);
    if (e[0] < e[1]) swapcols(V, e, 0, 1);
    m3<T> B = mmul(F, V);
    v3<T> b0 = {{B.a[0], B.a[3], B.a[6]}}, b1 = {{B.a[1], B.a[4], B.a[7]}}, b2 = {{B.a[2], B.a[5], B.a[8]}};
    T n0 = sqrt(vdot(b0, b0));
    v3<T> u0, u1, u2;
    const T tiny = sizeof(T) == 4 ? T(1e-30) : T(1e-300);
    if (n0 > tiny) {
        T inv = T(1) / n0;
        u0 = {{b0.a[0] * inv, b0.a[1] * inv, b0.a[2] * inv}};
    } else {
        u0 = {{T(1), T(0), T(0)}};
    }
    T d = vdot(u0, b1);
    v3<T> r1 = {{b1.a[0] - d * u0.a[0], b1.a[1] - d * u0.a[1], b1.a[2] - d * u0.a[2]}};
    T n1 = sqrt(vdot(r1, r1));
    if (n1 > tiny * (n0 + T(1))) {
        T inv = T(1) / n1;
        u1 = {{r1.a[0] * inv, r1.a[1] * inv, r1.a[2] * inv}};
    } else {
        // any unit vector orthogonal to u0
        v3<T> ax = fabs(u0.a[0]) < T(0.9) ? v3<T>{{T(1), T(0), T(0)}} : v3<T>{{T(0), T(1), T(0)}};
        T dd = vdot(u0, ax);
        v3<T> w = {{ax.a[0] - dd * u0.a[0], ax.a[1] - dd * u0.a[1], ax.a[2] - dd * u0.a[2]}};
        T inv = T(1) / sqrt(vdot(w, w));
        u1 = {{w.a[0] * inv, w.a[1] * inv, w.a[2] * inv}};
        n1 = T(0);
    }
    u2 = {{u0.a[1] * u1.a[2] - u0.a[2] * u1.a[1], u0.a[2] * u1.a[0] - u0.a[0] * u1.a[2],
           u0.a[0] * u1.a[1] - u0.a[1] * u1.a[0]}};
    sig.a[0] = n0;
    sig.a[1] = n1;
    sig.a[2] = vdot(u2, b2);
    U.a[0] = u0.a[0]; U.a[3] = u0.a[1]; U.a[6] = u0.a[2];
    U.a[1] = u1.a[0]; U.a[4] = u1.a[1]; U.a[7] = u1.a[2];
    U.a[2] = u2.a[0]; U.a[5] = u2.a[1]; U.a[8] = u2.a[2];
}

template <typename T>
DMPM_DI void polar3(const m3<T>& F, m3<T>& R, m3<T>& S) {
    m3<T> U, V;
    v3<T> s;
    svd3(F, U, s, V);
    R = mmult(U, V);
    m3<T> Vs;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Vs.a[3 * r + c] = V.a[3 * r + c] * s.a[c];
    S = mmult(Vs, V);
}

// ---------------------------------------------------------------------------
// Constitutive models.  MODEL 0: fixed-corotated (Elasticity.cpp:45-55),
// MODEL 1: Neo-Hookean.  Determinant, log J and F^{-T} in double.
template <int MODEL> struct Model;

template <> struct Model<1> {
    // P = mu F + (lam log J - mu) F^{-T}, J clamped >= 1e-6 inside the log.
    // det F, log J and F^{-T} are evaluated in FP64 (cancellation in det F
    // under strong compression); F^{-T} and the scalar factor are then
    // rounded to T and the products formed in T.
    template <typename T>
    static DMPM_DI m3<T> pk1(const m3<T>& F, T lam, T mu) {
        const m3<double> Fd = mcast<double>(F);
        const double J = mdet(Fd);
        const m3<T> Fit = mcast<T>(mscale(1.0 / J, mcof(Fd)));
        const T k = (T)((double)lam * log(fmax(J, 1e-6)) - (double)mu);
        m3<T> P;
#pragma unroll
        for (int i = 0; i < 9; ++i) P.a[i] = mu * F.a[i] + k * Fit.a[i];
        return P;
    }
    // Kirchhoff stress P F^T = mu F F^T + (lam log J - mu) I, since
    // F^{-T} F^T = I: what the P2G momentum and the adjoint's affine term use,
    // without the cofactor matrix and 1/J (det F and log J in FP64 as in pk1).
    template <typename T>
    static DMPM_DI m3<T> kirchhoff(const m3<T>& F, T lam, T mu) {
        const double J = mdet(mcast<double>(F));
        const T k = (T)((double)lam * log(fmax(J, 1e-6)) - (double)mu);
        m3<T> tau = mmult(F, F);
#pragma unroll
        for (int i = 0; i < 9; ++i) tau.a[i] *= mu;
        tau.a[0] += k;
        tau.a[4] += k;
        tau.a[8] += k;
        return tau;
    }
    // Gradient of <G, P(F) F^T> in F: G^T P + dP(F; G F), which for this
    // model reduces to mu (G + G^T) F + lam tr(G) F^{-T} (the log J terms
    // cancel; F^{-T} from FP64 det F and cofactors as in pk1).
    template <typename T>
    static DMPM_DI m3<T> kirchhoff_vjp(const m3<T>& F, T lam, T mu, const m3<T>& G) {
        const m3<double> Fd = mcast<double>(F);
        const double J = mdet(Fd);
        const m3<T> Fit = mcast<T>(mscale(1.0 / J, mcof(Fd)));
        m3<T> S;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) S.a[3 * r + c] = mu * (G.a[3 * r + c] + G.a[3 * c + r]);
        m3<T> R = mmul(S, F);
        const T c = lam * mtrace(G);
#pragma unroll
        for (int q = 0; q < 9; ++q) R.a[q] += c * Fit.a[q];
        return R;
    }
    // dP = mu dF + lam (F^{-T}:dF) F^{-T} + (mu - lam log J) F^{-T} dF^T F^{-T}
    // (FP64 as in pk1 for det F, log J and F^{-T}).
    template <typename T>
    static DMPM_DI m3<T> hvp(const m3<T>& F, T lam, T mu, const m3<T>& dF) {
        const m3<double> Fd = mcast<double>(F);
        const double J = mdet(Fd);
        const m3<T> Fit = mcast<T>(mscale(1.0 / J, mcof(Fd)));
        const T b = (T)((double)mu - (double)lam * log(fmax(J, 1e-6)));
        const T a = lam * mddot(Fit, dF);
        const m3<T> X = mmul(mmult(Fit, dF), Fit);  // F^{-T} dF^T F^{-T}
        m3<T> R;
#pragma unroll
        for (int i = 0; i < 9; ++i) R.a[i] = mu * dF.a[i] + a * Fit.a[i] + b * X.a[i];
        return R;
    }
};

template <> struct Model<0> {
    // The factors of F both the stress and its Hessian action use: the polar
    // decomposition F = R S, det F (raw and clamped to 1e-6) and F^{-T}, in
    // FP64.  The adjoint's stress gradient needs both, and forms them once.
    template <typename T>
    struct Parts {
        m3<T> R, S;
        m3<double> Fit;
        double Jr, J;
    };
    template <typename T>
    static DMPM_DI Parts<T> parts(const m3<T>& F) {
        m3<T> R, S;
        polar3(F, R, S);
        return parts(F, R, S);
    }
    // The same with the polar factors of F already known.
    template <typename T>
    static DMPM_DI Parts<T> parts(const m3<T>& F, const m3<T>& R, const m3<T>& S) {
        Parts<T> p;
        p.R = R;
        p.S = S;
        const m3<double> Fd = mcast<double>(F);
        p.Jr = mdet(Fd);
        p.J = fmax(p.Jr, 1e-6);
        p.Fit = mscale(1.0 / p.Jr, mcof(Fd));
        return p;
    }
    // Gradient of <G, P(F) F^T> in F from the parts of F.
    template <typename T>
    static DMPM_DI m3<T> kirchhoff_vjp(const Parts<T>& p, const m3<T>& F, T lam, T mu, const m3<T>& G) {
        return madd(mtmul(G, pk1(p, F, lam, mu)), hvp(p, lam, mu, mmul(G, F)));
    }
    // Kirchhoff stress P F^T, formed from P as the reference does.
    template <typename T>
    static DMPM_DI m3<T> kirchhoff(const m3<T>& F, T lam, T mu) {
        return mmult(pk1(F, lam, mu), F);
    }
    // Gradient of <G, P(F) F^T> in F: G^T P + dP(F; G F), with one polar
    // decomposition for both terms.
    template <typename T>
    static DMPM_DI m3<T> kirchhoff_vjp(const m3<T>& F, T lam, T mu, const m3<T>& G) {
        return kirchhoff_vjp(parts(F), F, lam, mu, G);
    }
    // P = 2 mu (F - R) + lam (J-1) J F^{-T}, J = max(det F, 1e-6)
    template <typename T>
    static DMPM_DI m3<T> pk1(const m3<T>& F, T lam, T mu) {
        return pk1(parts(F), F, lam, mu);
    }
    template <typename T>
    static DMPM_DI m3<T> pk1(const Parts<T>& p, const m3<T>& F, T lam, T mu) {
        double k = (double)lam * (p.J - 1.0) * p.J;
        m3<T> P;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            P.a[i] = (T)(2.0 * (double)mu * ((double)F.a[i] - (double)p.R.a[i]) + k * p.Fit.a[i]);
        return P;
    }
    // Analytic Hessian action of the reference (Elasticity.cpp:57-117):
    // 2mu dF + lam J F^{-T} (J F^{-T}:dF) + lam (J-1) Jraw [(F^{-T}:dF) F^{-T}
    // - F^{-T} dF^T F^{-T}] - 2mu R [w]_x with (tr S I - S) w = axial(R^T dF)
    // (+1e-4 I when |det| < 1e-6).
    template <typename T>
    static DMPM_DI m3<T> hvp(const m3<T>& F, T lam, T mu, const m3<T>& dF) {
        return hvp(parts(F), lam, mu, dF);
    }
    template <typename T>
    static DMPM_DI m3<T> hvp(const Parts<T>& p, T lam, T mu, const m3<T>& dF) {
        const m3<T>& R = p.R;
        const m3<T>& S = p.S;
        const m3<double>& Fit = p.Fit;
        const double Jr = p.Jr, J = p.J;
        m3<double> dFd = mcast<double>(dF);
        double fd = mddot(Fit, dFd);
        m3<double> X = mmul(mmult(Fit, dFd), Fit);
        double c1 = (double)lam * J * J * fd;
        double c2 = (double)lam * (J - 1.0) * Jr;
        m3<double> A;
        A.a[0] = (double)S.a[4] + S.a[8]; A.a[1] = -(double)S.a[1];      A.a[2] = -(double)S.a[2];
        A.a[3] = -(double)S.a[1];      A.a[4] = (double)S.a[0] + S.a[8]; A.a[5] = -(double)S.a[5];
        A.a[6] = -(double)S.a[2];      A.a[7] = -(double)S.a[5];      A.a[8] = (double)S.a[0] + S.a[4];
        double dA = mdet(A);
        if (fabs(dA) < 1e-6) {
            A.a[0] += 1e-4; A.a[4] += 1e-4; A.a[8] += 1e-4;
            dA = mdet(A);
        }
        m3<T> t = mtmul(R, dF);
        double b0 = (double)t.a[7] - (double)t.a[5];
        double b1 = (double)t.a[2] - (double)t.a[6];
        double b2 = (double)t.a[3] - (double)t.a[1];
        m3<double> Ac = mcof(A);  // A^{-1} = cof(A)^T / det
        double inv = 1.0 / dA;
        double w0 = (Ac.a[0] * b0 + Ac.a[3] * b1 + Ac.a[6] * b2) * inv;
        double w1 = (Ac.a[1] * b0 + Ac.a[4] * b1 + Ac.a[7] * b2) * inv;
        double w2 = (Ac.a[2] * b0 + Ac.a[5] * b1 + Ac.a[8] * b2) * inv;
        m3<double> W;
        W.a[0] = 0;   W.a[1] = -w2; W.a[2] = w1;
        W.a[3] = w2;  W.a[4] = 0;   W.a[5] = -w0;
        W.a[6] = -w1; W.a[7] = w0;  W.a[8] = 0;
        m3<double> RW = mmul(mcast<double>(R), W);
        m3<T> out;
#pragma unroll
        for (int i = 0; i < 9; ++i)
            out.a[i] = (T)(2.0 * (double)mu * dFd.a[i] + c1 * Fit.a[i] + c2 * (fd * Fit.a[i] - X.a[i]) - 2.0 * (double)mu * RW.a[i]);
        return out;
    }
};

}  // namespace dmpm
