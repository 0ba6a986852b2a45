// MLS-MPM step (P2G, grid update, G2P) and hand-written adjoints, 2D and 3D,
// FP32 and FP64.  Layouts: particle fields structure-of-arrays per robot,
// x/v [B][D][P], C/F [B][D*D][P] (row-major matrix entries); grids per robot,
// v [B][D][n^D], m [B][n^D], node index ((i*n + j)*n + k).
//
// P2G and the G2P adjoint scatter with global atomics; their summation order
// is the only run-to-run variation of the MPM path.

#include "mpm_core.hpp"

namespace dl {

// ----------------------------------------------------------- kernels
template <int D, typename T>
__global__ void k_p2g(MP<T> p, PView<D, T> pv, const T *act, const int32_t *act_id, int n_act,
                      T *F_out, T *gv, T *gm, int32_t *err) {
    const int k = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = pv.P;
    if (q >= P) return;
    T xs[D], vs[D];
    Mat<D, T> Cm, Fm;
    pv.load(k, q, xs, vs, Cm, Fm);
    const int id = act_id[(size_t)k * P + q];
    const T a = id >= 0 ? act[(size_t)k * n_act + id] : T(0);
    const Consti<D, T> c = constitutive<D, T>(Cm, Fm, a, p);
    store_mat<D, T>(F_out, k, q, P, c.Fn, false);
    const Stencil<D, T> s = stencil<D, T>(xs, p);
    if (!s.ok) {
        atomicOr(err, 1);
        return;
    }
    const int n = p.n;
    size_t G = 1;
    for (int d = 0; d < D; ++d) G *= n;
    T *gvk = gv + (size_t)k * D * G;
    T *gmk = gm + (size_t)k * G;
    constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
    for (int o = 0; o < K; ++o) {
        int off[D];
        unflat<D>(o, off);
        T W = T(1);
        T dpos[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            W *= s.w[off[d]][d];
            dpos[d] = (T(off[d]) - s.fx[d]) * p.dx;
        }
        const int node = node_of<D>(s.base, off, n);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            T m = p.p_mass * vs[i];
#pragma unroll
            for (int j = 0; j < D; ++j) m += c.affine.a[i][j] * dpos[j];
            atomic_add(gvk + (size_t)i * G + node, W * m);
        }
        atomic_add(gmk + node, W * p.p_mass);
    }
}

template <int D, typename T>
__device__ __forceinline__ bool on_boundary(int node, const MP<T> &p) {
    const int n = p.n;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        const int i = node % n;
        node /= n;
        if (i < p.bound || i > n - p.bound) return true;
    }
    return false;
}

template <int D, typename T>
__global__ void k_grid(MP<T> p, int G, const T *gv, const T *gm, T *gvo) {
    const int k = blockIdx.y, node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= G) return;
    const size_t bv = (size_t)k * D * G;
    const T inv_m = T(1) / (gm[(size_t)k * G + node] + T(1e-10));
    const bool bnd = on_boundary<D, T>(node, p);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        T vo = inv_m * gv[bv + (size_t)d * G + node];
        if (d == 1) vo += p.g_dt;
        gvo[bv + (size_t)d * G + node] = bnd ? T(0) : vo;
    }
}

template <int D, typename T>
__global__ void k_g2p(MP<T> p, int P, int G, const T *x, const T *gvo, T *x_out, T *v_out,
                      T *C_out) {
    const int k = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
    T xs[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xs[d] = x[bv + (size_t)d * P + q];
    const Stencil<D, T> s = stencil<D, T>(xs, p);
    if (!s.ok) return;
    const T *g = gvo + (size_t)k * D * G;
    T nv[D];
    T nC[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        nv[i] = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) nC[i][j] = T(0);
    }
    constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
    for (int o = 0; o < K; ++o) {
        int off[D];
        unflat<D>(o, off);
        T W = T(1), dp[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            W *= s.w[off[d]][d];
            dp[d] = T(off[d]) - s.fx[d];
        }
        const int node = node_of<D>(s.base, off, p.n);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const T gi = g[(size_t)i * G + node];
            nv[i] += W * gi;
#pragma unroll
            for (int j = 0; j < D; ++j) nC[i][j] += T(4) * p.inv_dx * W * gi * dp[j];
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v_out[bv + (size_t)d * P + q] = nv[d];
        x_out[bv + (size_t)d * P + q] = xs[d] + p.dt * nv[d];
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) C_out[bm + (size_t)(i * D + j) * P + q] = nC[i][j];
}

// G2P adjoint: (xb1, vb1, Cb1) = dL/d(x, v, C)[t+1] -> xb[t] (+=), gvo_b (atomic +=)
template <int D, typename T>
__global__ void k_g2p_adj(MP<T> p, int P, int G, const T *x, const T *gvo, const T *xb1,
                          const T *vb1, const T *Cb1, T *xb, T *gvo_b) {
    const int k = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
    T xs[D], vnb[D], Cb[D][D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        xs[d] = x[bv + (size_t)d * P + q];
        vnb[d] = vb1[bv + (size_t)d * P + q] + p.dt * xb1[bv + (size_t)d * P + q];
    }
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) Cb[i][j] = Cb1[bm + (size_t)(i * D + j) * P + q];
    const Stencil<D, T> s = stencil<D, T>(xs, p);
    T fxb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) fxb[d] = T(0);
    if (s.ok) {
        const T *g = gvo + (size_t)k * D * G;
        T *gb = gvo_b + (size_t)k * D * G;
        const T c4 = T(4) * p.inv_dx;
        constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
        for (int o = 0; o < K; ++o) {
            int off[D];
            unflat<D>(o, off);
            T W = T(1), dp[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                W *= s.w[off[d]][d];
                dp[d] = T(off[d]) - s.fx[d];
            }
            const int node = node_of<D>(s.base, off, p.n);
            T vo[D], Cdp[D];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                vo[i] = g[(size_t)i * G + node];
                T acc = T(0);
#pragma unroll
                for (int j = 0; j < D; ++j) acc += Cb[i][j] * dp[j];
                Cdp[i] = acc;
            }
            T Wb = T(0);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                Wb += vo[i] * (vnb[i] + c4 * Cdp[i]);
                atomic_add(gb + (size_t)i * G + node, W * (vnb[i] + c4 * Cdp[i]));
            }
#pragma unroll
            for (int j = 0; j < D; ++j) {
                T CtV = T(0);
#pragma unroll
                for (int i = 0; i < D; ++i) CtV += Cb[i][j] * vo[i];
                fxb[j] -= c4 * W * CtV;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T dWd = s.dw[off[d]][d];
#pragma unroll
                for (int e = 0; e < D; ++e)
                    if (e != d) dWd *= s.w[off[e]][e];
                fxb[d] += Wb * dWd;
            }
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        xb[bv + (size_t)d * P + q] = xb1[bv + (size_t)d * P + q] + p.inv_dx * fxb[d];
}

template <int D, typename T>
__global__ void k_grid_adj(MP<T> p, int G, const T *gv, const T *gm, const T *gvo_b, T *gv_b,
                           T *gm_b) {
    const int k = blockIdx.y, node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= G) return;
    const size_t bv = (size_t)k * D * G;
    const bool bnd = on_boundary<D, T>(node, p);
    const T m = gm[(size_t)k * G + node] + T(1e-10);
    T mb = T(0);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T ob = bnd ? T(0) : gvo_b[bv + (size_t)d * G + node];
        gv_b[bv + (size_t)d * G + node] = ob / m;
        mb -= ob * gv[bv + (size_t)d * G + node] / (m * m);
    }
    gm_b[(size_t)k * G + node] = mb;
}

// P2G adjoint: grid adjoints (gv_b, gm_b) and dL/dF[t+1] -> xb (+=), vb, Cb, Fb
// (written), act_b (atomic +=).
template <int D, typename T>
__global__ void k_p2g_adj(MP<T> p, PView<D, T> pv, const T *act, const int32_t *act_id,
                          int n_act, int G, const T *gv_b, const T *gm_b, const T *Fb1, T *xb,
                          T *vb, T *Cb, T *Fb, T *act_b) {
    const int k = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    const int P = pv.P;
    if (q >= P) return;
    T xs[D], vs[D];
    Mat<D, T> Cm, Fm;
    pv.load(k, q, xs, vs, Cm, Fm);
    const int id = act_id[(size_t)k * P + q];
    const T a = id >= 0 ? act[(size_t)k * n_act + id] : T(0);
    const Consti<D, T> c = constitutive<D, T>(Cm, Fm, a, p);
    const Stencil<D, T> s = stencil<D, T>(xs, p);
    const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
    Mat<D, T> affb, Fnb;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            affb.a[i][j] = T(0);
            Fnb.a[i][j] = Fb1 ? Fb1[bm + (size_t)(i * D + j) * P + q] : T(0);
        }
    T vacc[D], fxb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) vacc[d] = fxb[d] = T(0);
    if (s.ok) {
        const T *gb = gv_b + (size_t)k * D * G;
        const T *mbk = gm_b + (size_t)k * G;
        constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
        for (int o = 0; o < K; ++o) {
            int off[D];
            unflat<D>(o, off);
            T W = T(1), dpos[D];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                W *= s.w[off[d]][d];
                dpos[d] = (T(off[d]) - s.fx[d]) * p.dx;
            }
            const int node = node_of<D>(s.base, off, p.n);
            T g[D];
#pragma unroll
            for (int i = 0; i < D; ++i) g[i] = gb[(size_t)i * G + node];
            const T h = mbk[node];
            T Wb = h * p.p_mass;
#pragma unroll
            for (int i = 0; i < D; ++i) {
                T cv = p.p_mass * vs[i];
#pragma unroll
                for (int j = 0; j < D; ++j) cv += c.affine.a[i][j] * dpos[j];
                Wb += g[i] * cv;
                vacc[i] += W * p.p_mass * g[i];
#pragma unroll
                for (int j = 0; j < D; ++j) affb.a[i][j] += W * g[i] * dpos[j];
            }
#pragma unroll
            for (int j = 0; j < D; ++j) {
                T atg = T(0);
#pragma unroll
                for (int i = 0; i < D; ++i) atg += c.affine.a[i][j] * g[i];
                fxb[j] -= p.dx * W * atg;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T dWd = s.dw[off[d]][d];
#pragma unroll
                for (int e = 0; e < D; ++e)
                    if (e != d) dWd *= s.w[off[e]][e];
                fxb[d] += Wb * dWd;
            }
        }
    }
    Mat<D, T> Fbm, Cbm;
    const T abar = consti_adj<D, T>(p, c, Cm, Fm, affb, Fnb, Fbm, Cbm);
    if (id >= 0) atomic_add(act_b + (size_t)k * n_act + id, abar);
    store_mat<D, T>(Fb, k, q, P, Fbm, false);
    store_mat<D, T>(Cb, k, q, P, Cbm, false);
#pragma unroll
    for (int d = 0; d < D; ++d) {
        vb[bv + (size_t)d * P + q] = vacc[d];
        xb[bv + (size_t)d * P + q] += p.inv_dx * fxb[d];
    }
}

template <typename T>
__global__ void k_fill(T *p, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Sensor vector of an MPM robot: [sin | vcom (3) | per-group mean offset from
// the COM / scale (3 per group) | tv / max_speed (3) | th].  Block per robot.
template <typename T>
__global__ void k_mpm_sensor(int P, int Gs, int n_sin, int run_period, int t, T scale, T max_speed,
                             const T *x, const T *v, const int32_t *group, const T *tgt,
                             const T *inv_count, T *in, T *center, T *vcom) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *acc = reinterpret_cast<T *>(smem_raw);      // [3*Gs + 6]
    T *red = acc + 3 * Gs + 6;
    const int k = blockIdx.x;
    const int I = n_sin + 3 + 3 * Gs + 4;
    for (int i = threadIdx.x; i < 3 * Gs + 6; i += blockDim.x) acc[i] = T(0);
    __syncthreads();
    const size_t bv = (size_t)k * 3 * P;
    T s[6] = {0, 0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < P; q += blockDim.x) {
        const int g = group[(size_t)k * P + q];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T xv = x[bv + (size_t)d * P + q];
            s[d] += xv;
            s[3 + d] += v[bv + (size_t)d * P + q];
            atomicAdd(acc + 3 * g + d, xv);
        }
    }
    block_sum<T, 6>(s, red);
    __syncthreads();
    const T invP = T(1) / T(P);
    T c[3] = {s[0] * invP, s[1] * invP, s[2] * invP};
    T vc[3] = {s[3] * invP, s[4] * invP, s[5] * invP};
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) {
            if (center) center[k * 3 + d] = c[d];
            if (vcom) vcom[k * 3 + d] = vc[d];
        }
    }
    T *o = in + (size_t)k * I;
    const T *tg = tgt + (size_t)k * 5;
    for (int j = threadIdx.x; j < I; j += blockDim.x) {
        T val;
        if (j < n_sin) {
            val = T(::sin(2.0 * M_PI / run_period * (double)t + 2.0 * M_PI / n_sin * (double)j));
        } else if (j < n_sin + 3) {
            val = vc[j - n_sin];
        } else if (j < n_sin + 3 + 3 * Gs) {
            const int q = j - n_sin - 3, g = q / 3, d = q - 3 * g;
            val = (acc[q] * inv_count[(size_t)k * Gs + g] - c[d]) / scale;
        } else if (j < n_sin + 6 + 3 * Gs) {
            val = tg[j - n_sin - 3 - 3 * Gs] / max_speed;
        } else {
            val = tg[3];
        }
        o[j] = val;
    }
}

template <typename T>
__global__ void k_mpm_sensor_adj(int P, int Gs, int n_sin, T scale, const int32_t *group,
                                 const T *inv_count, const T *in_b, T *xb, T *vb) {
    const int k = blockIdx.y, q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const int I = n_sin + 3 + 3 * Gs + 4;
    const T *g = in_b + (size_t)k * I;
    const int grp = group[(size_t)k * P + q];
    const T invP = T(1) / T(P);
    const size_t bv = (size_t)k * 3 * P;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        T sum = T(0);
        for (int h = 0; h < Gs; ++h) sum += g[n_sin + 3 + 3 * h + d];
        const T gx = (g[n_sin + 3 + 3 * grp + d] * inv_count[(size_t)k * Gs + grp] - sum * invP) / scale;
        xb[bv + (size_t)d * P + q] += gx;
        vb[bv + (size_t)d * P + q] += g[n_sin + d] * invP;
    }
}

}  // namespace dl

// ============================================================== C ABI
using namespace dl;

static inline int gblk(long n, int b) { return (int)((n + b - 1) / b); }

static int check_mp(const dl_mpm_params *c) {
    DL_CHECK_ARG(c != nullptr, "MPM params NULL");
    DL_CHECK_ARG(c->dim == 2 || c->dim == 3, "MPM: dim must be 2 or 3");
    DL_CHECK_ARG(c->n_grid >= 8 && c->n_grid <= 1024, "MPM: n_grid out of range");
    DL_CHECK_ARG(c->act_axis >= 0 && c->act_axis < c->dim, "MPM: bad actuation axis");
    return 0;
}

static long grid_nodes(const dl_mpm_params *c) {
    long G = 1;
    for (int d = 0; d < c->dim; ++d) G *= c->n_grid;
    return G;
}

#define MPM_DISPATCH(c, dtype, ...)                                                   \
    do {                                                                              \
        if ((c)->dim == 2 && (dtype) == DL_F32) {                                     \
            constexpr int D = 2; using real = float; __VA_ARGS__;                     \
        } else if ((c)->dim == 2 && (dtype) == DL_F64) {                              \
            constexpr int D = 2; using real = double; __VA_ARGS__;                    \
        } else if ((c)->dim == 3 && (dtype) == DL_F32) {                              \
            constexpr int D = 3; using real = float; __VA_ARGS__;                     \
        } else if ((c)->dim == 3 && (dtype) == DL_F64) {                              \
            constexpr int D = 3; using real = double; __VA_ARGS__;                    \
        } else {                                                                      \
            dl::set_error("MPM: unsupported dtype %d", (int)(dtype));                 \
            return DL_E_ARG;                                                          \
        }                                                                             \
    } while (0)

extern "C" {

int32_t dl_mpm_fill(int32_t dtype, void *ptr, int64_t n, double value, void *stream) {
    DL_CHECK_ARG(ptr && n >= 0, "dl_mpm_fill: bad arguments");
    if (n == 0) return 0;
    const int nb = (int)std::min<int64_t>((n + 255) / 256, 4096);
    if (dtype == DL_F32)
        hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                           (float *)ptr, (size_t)n, (float)value);
    else if (dtype == DL_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                           (double *)ptr, (size_t)n, value);
    else
        DL_CHECK_ARG(false, "dl_mpm_fill: bad dtype");
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_p2g(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P, const void *x,
                   const void *v, const void *C, const void *F, const void *act,
                   const int32_t *act_id, int32_t n_act, void *F_out, void *grid_v,
                   void *grid_m, int32_t *err, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && P > 0 && n_act > 0 && x && v && C && F && act && act_id && F_out &&
                     grid_v && grid_m && err,
                 "dl_mpm_p2g: bad arguments");
    MPM_DISPATCH(c, dtype, {
        PView<D, real> pv{(const real *)x, (const real *)v, (const real *)C, (const real *)F, P};
        hipLaunchKernelGGL((k_p2g<D, real>), dim3(gblk(P, 256), B), dim3(256), 0,
                           (hipStream_t)stream, make_mp<real>(*c), pv, (const real *)act, act_id,
                           n_act, (real *)F_out, (real *)grid_v, (real *)grid_m, err);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_grid(int32_t dtype, const dl_mpm_params *params, int32_t B, const void *grid_v,
                    const void *grid_m, void *grid_v_out, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && grid_v && grid_m && grid_v_out, "dl_mpm_grid: bad arguments");
    const long G = grid_nodes(c);
    MPM_DISPATCH(c, dtype, hipLaunchKernelGGL((k_grid<D, real>), dim3(gblk(G, 256), B), dim3(256),
                                              0, (hipStream_t)stream, make_mp<real>(*c), (int)G,
                                              (const real *)grid_v, (const real *)grid_m,
                                              (real *)grid_v_out));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_g2p(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P, const void *x,
                   const void *grid_v_out, void *x_out, void *v_out, void *C_out, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && P > 0 && x && grid_v_out && x_out && v_out && C_out,
                 "dl_mpm_g2p: bad arguments");
    const long G = grid_nodes(c);
    MPM_DISPATCH(c, dtype, hipLaunchKernelGGL((k_g2p<D, real>), dim3(gblk(P, 256), B), dim3(256),
                                              0, (hipStream_t)stream, make_mp<real>(*c), P, (int)G,
                                              (const real *)x, (const real *)grid_v_out,
                                              (real *)x_out, (real *)v_out, (real *)C_out));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_g2p_adj(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P, const void *x,
                       const void *grid_v_out, const void *x_grad1, const void *v_grad1,
                       const void *C_grad1, void *x_grad, void *grid_v_out_grad, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && P > 0 && x && grid_v_out && x_grad1 && v_grad1 && C_grad1 && x_grad &&
                     grid_v_out_grad,
                 "dl_mpm_g2p_adj: bad arguments");
    const long G = grid_nodes(c);
    MPM_DISPATCH(c, dtype, hipLaunchKernelGGL((k_g2p_adj<D, real>), dim3(gblk(P, 256), B),
                                              dim3(256), 0, (hipStream_t)stream, make_mp<real>(*c),
                                              P, (int)G, (const real *)x, (const real *)grid_v_out,
                                              (const real *)x_grad1, (const real *)v_grad1,
                                              (const real *)C_grad1, (real *)x_grad,
                                              (real *)grid_v_out_grad));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_grid_adj(int32_t dtype, const dl_mpm_params *params, int32_t B, const void *grid_v,
                        const void *grid_m, const void *grid_v_out_grad, void *grid_v_grad,
                        void *grid_m_grad, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && grid_v && grid_m && grid_v_out_grad && grid_v_grad && grid_m_grad,
                 "dl_mpm_grid_adj: bad arguments");
    const long G = grid_nodes(c);
    MPM_DISPATCH(c, dtype, hipLaunchKernelGGL((k_grid_adj<D, real>), dim3(gblk(G, 256), B),
                                              dim3(256), 0, (hipStream_t)stream, make_mp<real>(*c),
                                              (int)G, (const real *)grid_v, (const real *)grid_m,
                                              (const real *)grid_v_out_grad, (real *)grid_v_grad,
                                              (real *)grid_m_grad));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_p2g_adj(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P, const void *x,
                       const void *v, const void *C, const void *F, const void *act,
                       const int32_t *act_id, int32_t n_act, const void *grid_v_grad,
                       const void *grid_m_grad, const void *F_grad1, void *x_grad, void *v_grad,
                       void *C_grad, void *F_grad, void *act_grad, void *stream) {
    const dl_mpm_params *c = params;
    if (int e = check_mp(c)) return e;
    DL_CHECK_ARG(B > 0 && P > 0 && n_act > 0 && x && v && C && F && act && act_id &&
                     grid_v_grad && grid_m_grad && x_grad && v_grad && C_grad && F_grad &&
                     act_grad,
                 "dl_mpm_p2g_adj: bad arguments");
    const long G = grid_nodes(c);
    MPM_DISPATCH(c, dtype, {
        PView<D, real> pv{(const real *)x, (const real *)v, (const real *)C, (const real *)F, P};
        hipLaunchKernelGGL((k_p2g_adj<D, real>), dim3(gblk(P, 256), B), dim3(256), 0,
                           (hipStream_t)stream, make_mp<real>(*c), pv, (const real *)act, act_id,
                           n_act, (int)G, (const real *)grid_v_grad, (const real *)grid_m_grad,
                           (const real *)F_grad1, (real *)x_grad, (real *)v_grad, (real *)C_grad,
                           (real *)F_grad, (real *)act_grad);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_sensor(int32_t dtype, int32_t B, int32_t P, int32_t n_groups, int32_t n_sin,
                      int32_t run_period, int32_t t, double offset_scale, double max_speed,
                      const void *x, const void *v, const int32_t *group, const void *targets,
                      const void *inv_count, void *in, void *center, void *vcom, void *stream) {
    DL_CHECK_ARG(B > 0 && P > 0 && n_groups > 0 && n_sin >= 0 && run_period > 0 && x && v &&
                     group && targets && inv_count && in,
                 "dl_mpm_sensor: bad arguments");
    const size_t sh_elems = 3 * (size_t)n_groups + 6 + 24;
    if (dtype == DL_F32)
        hipLaunchKernelGGL(k_mpm_sensor<float>, dim3(B), dim3(256), sh_elems * sizeof(float),
                           (hipStream_t)stream, P, n_groups, n_sin, run_period, t,
                           (float)offset_scale, (float)max_speed, (const float *)x,
                           (const float *)v, group, (const float *)targets,
                           (const float *)inv_count, (float *)in, (float *)center, (float *)vcom);
    else if (dtype == DL_F64)
        hipLaunchKernelGGL(k_mpm_sensor<double>, dim3(B), dim3(256), sh_elems * sizeof(double),
                           (hipStream_t)stream, P, n_groups, n_sin, run_period, t, offset_scale,
                           max_speed, (const double *)x, (const double *)v, group,
                           (const double *)targets, (const double *)inv_count, (double *)in,
                           (double *)center, (double *)vcom);
    else
        DL_CHECK_ARG(false, "dl_mpm_sensor: bad dtype");
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mpm_sensor_adj(int32_t dtype, int32_t B, int32_t P, int32_t n_groups, int32_t n_sin,
                          double offset_scale, const int32_t *group, const void *inv_count,
                          const void *in_grad, void *x_grad, void *v_grad, void *stream) {
    DL_CHECK_ARG(B > 0 && P > 0 && n_groups > 0 && group && inv_count && in_grad && x_grad &&
                     v_grad,
                 "dl_mpm_sensor_adj: bad arguments");
    if (dtype == DL_F32)
        hipLaunchKernelGGL(k_mpm_sensor_adj<float>, dim3(gblk(P, 256), B), dim3(256), 0,
                           (hipStream_t)stream, P, n_groups, n_sin, (float)offset_scale, group,
                           (const float *)inv_count, (const float *)in_grad, (float *)x_grad,
                           (float *)v_grad);
    else if (dtype == DL_F64)
        hipLaunchKernelGGL(k_mpm_sensor_adj<double>, dim3(gblk(P, 256), B), dim3(256), 0,
                           (hipStream_t)stream, P, n_groups, n_sin, offset_scale, group,
                           (const double *)inv_count, (const double *)in_grad, (double *)x_grad,
                           (double *)v_grad);
    else
        DL_CHECK_ARG(false, "dl_mpm_sensor_adj: bad dtype");
    DL_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
