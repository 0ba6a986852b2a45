// Shared device code of the MLS-MPM kernels: parameters, small matrices,
// polar decomposition and its adjoint, the quadratic B-spline stencil, the
// particle loader and the fixed-corotated constitutive update with its
// adjoint.  Follows multitask/solver_mpm.py:83-165.
#pragma once

#include "dl_common.hpp"

namespace dl {

template <typename T> struct MP {
    int n, bound, act_axis;
    T dx, inv_dx, dt, g_dt, p_vol, p_mass, mu, la, act_strength, kappa;
    // uneven batches: particles [0, np[k]) of robot k are its own (the rest of
    // its P slots are padding, never read); robot k advances steps t < hz[k]
    // only (t = the step a launch computes).  nullptr: every robot has P
    // particles and the whole horizon.
    const int32_t *np = nullptr, *hz = nullptr;
    int t = 0;
};

// Robot k does not take step t (past its horizon).
template <typename T> __host__ __device__ __forceinline__ bool idle_at(const MP<T> &p, int k, int t) {
    return p.hz && t >= p.hz[k];
}
// Particles of robot k (P when the batch is even).
template <typename T> __host__ __device__ __forceinline__ int n_part(const MP<T> &p, int k, int P) {
    return p.np ? min(p.np[k], P) : P;
}

template <typename T>
static MP<T> make_mp(const dl_mpm_params &c) {
    MP<T> p;
    p.n = c.n_grid; p.bound = c.bound; p.act_axis = c.act_axis;
    p.dx = T(1.0 / c.n_grid); p.inv_dx = T(c.n_grid);
    p.dt = T(c.dt); p.g_dt = T(c.dt * c.gravity); p.p_vol = T(c.p_vol); p.p_mass = T(c.p_mass);
    p.mu = T(c.mu); p.la = T(c.la); p.act_strength = T(c.act_strength);
    p.kappa = T(c.dt * c.p_vol * 4.0 * c.n_grid * (double)c.n_grid);
    return p;
}

// ----------------------------------------------------------- small matrices
template <int D, typename T> struct Mat { T a[D][D]; };

template <int D, typename T>
__device__ __forceinline__ Mat<D, T> mul(const Mat<D, T> &A, const Mat<D, T> &B) {
    Mat<D, T> C;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < D; ++k) s += A.a[i][k] * B.a[k][j];
            C.a[i][j] = s;
        }
    return C;
}

template <int D, typename T>
__device__ __forceinline__ Mat<D, T> transpose(const Mat<D, T> &A) {
    Mat<D, T> C;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) C.a[i][j] = A.a[j][i];
    return C;
}

template <typename T>
__device__ __forceinline__ T det(const Mat<2, T> &A) { return A.a[0][0] * A.a[1][1] - A.a[0][1] * A.a[1][0]; }
template <typename T>
__device__ __forceinline__ T det(const Mat<3, T> &A) {
    return A.a[0][0] * (A.a[1][1] * A.a[2][2] - A.a[1][2] * A.a[2][1]) -
           A.a[0][1] * (A.a[1][0] * A.a[2][2] - A.a[1][2] * A.a[2][0]) +
           A.a[0][2] * (A.a[1][0] * A.a[2][1] - A.a[1][1] * A.a[2][0]);
}

// cofactor matrix (d det / dA)
template <typename T>
__device__ __forceinline__ Mat<2, T> cof(const Mat<2, T> &A) {
    Mat<2, T> C;
    C.a[0][0] = A.a[1][1]; C.a[0][1] = -A.a[1][0];
    C.a[1][0] = -A.a[0][1]; C.a[1][1] = A.a[0][0];
    return C;
}
template <typename T>
__device__ __forceinline__ Mat<3, T> cof(const Mat<3, T> &A) {
    Mat<3, T> C;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            C.a[i][j] = A.a[i1][j1] * A.a[i2][j2] - A.a[i1][j2] * A.a[i2][j1];
        }
    return C;
}

// Rotation factor of the polar decomposition A = R S.
// 2D: closed form (rotation angle from the trace/skew parts).
template <typename T>
__device__ __forceinline__ Mat<2, T> polar_R(const Mat<2, T> &A) {
    const T x = A.a[0][0] + A.a[1][1], y = A.a[1][0] - A.a[0][1];
    const T sc = T(1) / M<T>::sqrt(x * x + y * y);
    const T c = x * sc, s = y * sc;
    Mat<2, T> R;
    R.a[0][0] = c; R.a[0][1] = -s; R.a[1][0] = s; R.a[1][1] = c;
    return R;
}
// 3D: scaled Newton iteration X <- (g X + X^{-T} / g) / 2, g = |det X|^(-1/3)
// (Higham), converged to working precision for det A > 0.
template <typename T>
__device__ __forceinline__ Mat<3, T> polar_R(const Mat<3, T> &A) {
    // The scaling g only changes the speed of convergence, not the limit.
    // Stop once an update moves the nine entries by less than a few ulps in
    // total (quadratic convergence: the iterate is then exact to rounding).
    Mat<3, T> X = A;
    const T tol = sizeof(T) == 4 ? T(4e-6) : T(1e-14);
    for (int it = 0; it < 30; ++it) {
        const T dt_ = det(X);
        const Mat<3, T> Cf = cof(X);          // X^{-T} = cof(X) / det
        const T g = (it < 6) ? M<T>::rcbrt(fabs(dt_)) : T(1);
        const T h = T(0.5) / (g * dt_);
        T diff = T(0);
        Mat<3, T> Y;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                Y.a[i][j] = T(0.5) * g * X.a[i][j] + h * Cf.a[i][j];
                diff += fabs(Y.a[i][j] - X.a[i][j]);
            }
        X = Y;
        if (it >= 2 && diff < tol) break;
    }
    return X;
}

// Adjoint of R = polar(A): Abar = 2 R Y, Y skew with Y S + S Y = skew(R^T Rbar).
template <typename T>
__device__ __forceinline__ Mat<2, T> polar_adj(const Mat<2, T> &A, const Mat<2, T> &R,
                                               const Mat<2, T> &Rb) {
    const Mat<2, T> S = mul(transpose(R), A);
    const Mat<2, T> G = mul(transpose(R), Rb);
    const T k = T(0.5) * (G.a[1][0] - G.a[0][1]);
    const T y = k / (S.a[0][0] + S.a[1][1]);
    Mat<2, T> Y;
    Y.a[0][0] = T(0); Y.a[0][1] = -y; Y.a[1][0] = y; Y.a[1][1] = T(0);
    Mat<2, T> out = mul(R, Y);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) out.a[i][j] *= T(2);
    return out;
}
template <typename T>
__device__ __forceinline__ Mat<3, T> polar_adj(const Mat<3, T> &A, const Mat<3, T> &R,
                                               const Mat<3, T> &Rb) {
    Mat<3, T> S = mul(transpose(R), A);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i + 1; j < 3; ++j) S.a[i][j] = S.a[j][i] = T(0.5) * (S.a[i][j] + S.a[j][i]);
    const Mat<3, T> G = mul(transpose(R), Rb);
    const T k[3] = {T(0.5) * (G.a[2][1] - G.a[1][2]), T(0.5) * (G.a[0][2] - G.a[2][0]),
                    T(0.5) * (G.a[1][0] - G.a[0][1])};
    // solve (tr S I - S) w = k
    const T tr = S.a[0][0] + S.a[1][1] + S.a[2][2];
    Mat<3, T> Mm;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Mm.a[i][j] = (i == j ? tr : T(0)) - S.a[i][j];
    const Mat<3, T> Cf = cof(Mm);
    const T dm = det(Mm);
    T w[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) w[i] = (Cf.a[0][i] * k[0] + Cf.a[1][i] * k[1] + Cf.a[2][i] * k[2]) / dm;
    Mat<3, T> Y;
    Y.a[0][0] = T(0); Y.a[0][1] = -w[2]; Y.a[0][2] = w[1];
    Y.a[1][0] = w[2]; Y.a[1][1] = T(0); Y.a[1][2] = -w[0];
    Y.a[2][0] = -w[1]; Y.a[2][1] = w[0]; Y.a[2][2] = T(0);
    Mat<3, T> out = mul(R, Y);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out.a[i][j] *= T(2);
    return out;
}

// ----------------------------------------------------------- stencil
template <int D, typename T>
struct Stencil {
    int base[D];
    T fx[D];
    T w[3][D], dw[3][D];
    bool ok;        // the 3^D stencil lies inside the grid
};

// base = trunc(x/dx - 0.5) like the reference's integer cast; the stencil
// base..base+2 is inside the grid iff 0 <= base <= n-3, i.e. -1 < x/dx - 0.5
// < n-2.  That test is made on the real value, before the conversion, so
// that NaN, infinities and coordinates beyond the int range come out as
// outside (their base is then set to 0 and must not be used).
template <int D, typename T>
__device__ __forceinline__ Stencil<D, T> stencil(const T *x, const MP<T> &p) {
    Stencil<D, T> s;
    s.ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T y = x[d] * p.inv_dx - T(0.5);
        const bool in = y > T(-1) && y < T(p.n - 2);
        s.ok = s.ok && in;
        s.base[d] = in ? (int)y : 0;
        const T f = y + T(0.5) - T(s.base[d]);
        s.fx[d] = f;
        s.w[0][d] = T(0.5) * (T(1.5) - f) * (T(1.5) - f);
        s.w[1][d] = T(0.75) - (f - T(1)) * (f - T(1));
        s.w[2][d] = T(0.5) * (f - T(0.5)) * (f - T(0.5));
        s.dw[0][d] = f - T(1.5);
        s.dw[1][d] = T(-2) * (f - T(1));
        s.dw[2][d] = f - T(0.5);
    }
    return s;
}

template <int D>
__device__ __forceinline__ void unflat(int o, int off[D]) {
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        off[d] = o % 3;
        o /= 3;
    }
}

template <int D>
__device__ __forceinline__ int node_of(const int base[D], const int off[D], int n) {
    int id = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) id = id * n + base[d] + off[d];
    return id;
}


template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) { atomicAdd(p, v); }

// ----------------------------------------------------------- particle io
template <int D, typename T>
struct PView {
    const T *x, *v, *C, *F;
    int P;
    __device__ void load(int k, int q, T xs[D], T vs[D], Mat<D, T> &Cm, Mat<D, T> &Fm) const {
        const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xs[d] = x[bv + (size_t)d * P + q];
            vs[d] = v[bv + (size_t)d * P + q];
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) {
                Cm.a[i][j] = C[bm + (size_t)(i * D + j) * P + q];
                Fm.a[i][j] = F[bm + (size_t)(i * D + j) * P + q];
            }
    }
};

template <int D, typename T>
__device__ __forceinline__ void store_mat(T *dst, int k, int q, int P, const Mat<D, T> &A, bool add) {
    const size_t bm = (size_t)k * D * D * P;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            T &r = dst[bm + (size_t)(i * D + j) * P + q];
            r = add ? r + A.a[i][j] : A.a[i][j];
        }
}

// One-hot weights of axis e (a runtime index).  Products with them keep
// matrices in registers, where an indexed access A[i][e] would move A to
// scratch memory.
template <int D, typename T>
__device__ __forceinline__ void onehot(int e, T w[D]) {
#pragma unroll
    for (int j = 0; j < D; ++j) w[j] = (e == j) ? T(1) : T(0);
}

// Column e of A.
template <int D, typename T>
__device__ __forceinline__ void column(const Mat<D, T> &A, int e, T col[D]) {
    T w[D];
    onehot<D, T>(e, w);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T v = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) v += A.a[i][j] * w[j];
        col[i] = v;
    }
}

// Deformation gradient of the next state: (I + dt C) F (solver_mpm.py:89, 95).
template <int D, typename T>
__device__ __forceinline__ Mat<D, T> f_next(const Mat<D, T> &Cm, const Mat<D, T> &Fm,
                                            const MP<T> &p) {
    Mat<D, T> Mx;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Mx.a[i][j] = (i == j ? T(1) : T(0)) + p.dt * Cm.a[i][j];
    return mul(Mx, Fm);
}

// Forward constitutive quantities of one particle.
template <int D, typename T>
struct Consti {
    Mat<D, T> Fn, R, sigma, affine;
    T J, a;
};

template <int D, typename T>
__device__ __forceinline__ Consti<D, T> constitutive(const Mat<D, T> &Cm, const Mat<D, T> &Fm,
                                                     T a, const MP<T> &p) {
    Consti<D, T> q;
    q.Fn = f_next<D, T>(Cm, Fm, p);
    q.J = det(q.Fn);
    q.R = polar_R(q.Fn);
    q.a = a;
    Mat<D, T> Pm;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Pm.a[i][j] = q.Fn.a[i][j] - q.R.a[i][j];
    const Mat<D, T> FnT = transpose(q.Fn);
    const Mat<D, T> s1 = mul(Pm, FnT);
    const T s2 = p.la * (q.J - T(1)) * q.J;
    const T sa = p.act_strength * a;
    T fe[D];
    column<D, T>(q.Fn, p.act_axis, fe);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            // F A F^T with A = sa e e^T -> sa F[i][e] F[j][e]
            q.sigma.a[i][j] = T(2) * p.mu * s1.a[i][j] + (i == j ? s2 : T(0)) +
                              sa * fe[i] * fe[j];
            q.affine.a[i][j] = -p.kappa * q.sigma.a[i][j] + p.p_mass * Cm.a[i][j];
        }
    return q;
}

// Adjoint of the P2G constitutive update of one particle.  Inputs: the
// forward quantities c (from Cm, Fm and the actuation), affb = dL/d(affine)
// gathered from the grid, Fnb = dL/dF[t+1] from the next step.  Outputs:
// Fbm_out = dL/dF[t], Cbm_out = dL/dC[t]; returns dL/da for the particle's
// actuation (the caller drops it for passive particles).
template <int D, typename T>
__device__ __forceinline__ T consti_adj(const MP<T> &p, const Consti<D, T> &c, const Mat<D, T> &Cm,
                                        const Mat<D, T> &Fm, const Mat<D, T> &affb,
                                        Mat<D, T> Fnb, Mat<D, T> &Fbm_out, Mat<D, T> &Cbm_out) {
    T act_bar = T(0);
    // affine = -kappa sigma + m C
    Mat<D, T> Cbm, sb;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            Cbm.a[i][j] = p.p_mass * affb.a[i][j];
            sb.a[i][j] = -p.kappa * affb.a[i][j];
        }
    // sigma = 2 mu (Fn - R) Fn^T + la (J-1) J I + sa Fn e e^T Fn^T
    Mat<D, T> Pm;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Pm.a[i][j] = c.Fn.a[i][j] - c.R.a[i][j];
    const Mat<D, T> Pb = mul(sb, c.Fn);               // / (2 mu)
    const Mat<D, T> sTP = mul(transpose(sb), Pm);
    T trs = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) trs += sb.a[i][i];
    const T Jb = p.la * (T(2) * c.J - T(1)) * trs;
    const Mat<D, T> cf = cof(c.Fn);
    const int e = p.act_axis;
    const T sa = p.act_strength * c.a;
    Mat<D, T> Rb;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            Fnb.a[i][j] += T(2) * p.mu * (sTP.a[i][j] + Pb.a[i][j]) + Jb * cf.a[i][j];
            Rb.a[i][j] = -T(2) * p.mu * Pb.a[i][j];
        }
    // (sb + sb^T) Fn A : column e only
    T fe[D], we[D];
    column<D, T>(c.Fn, e, fe);
    onehot<D, T>(e, we);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) acc += (sb.a[i][j] + sb.a[j][i]) * fe[j];
#pragma unroll
        for (int j = 0; j < D; ++j) Fnb.a[i][j] += we[j] * (sa * acc);
    }
    {
        T ab = T(0);
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) ab += fe[i] * sb.a[i][j] * fe[j];
        act_bar = p.act_strength * ab;
    }
    const Mat<D, T> pa = polar_adj(c.Fn, c.R, Rb);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Fnb.a[i][j] += pa.a[i][j];
    // Fn = (I + dt C) F
    Mat<D, T> Mx;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Mx.a[i][j] = (i == j ? T(1) : T(0)) + p.dt * Cm.a[i][j];
    Fbm_out = mul(transpose(Mx), Fnb);
    const Mat<D, T> FnbFt = mul(Fnb, transpose(Fm));
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Cbm.a[i][j] += p.dt * FnbFt.a[i][j];
    Cbm_out = Cbm;
    return act_bar;
}

}  // namespace dl
