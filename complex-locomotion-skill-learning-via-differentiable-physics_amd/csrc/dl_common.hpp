// Shared device code for the diffloco kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>

#include <mutex>
#include <unordered_map>

#include "diffloco.h"

namespace dl {

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...);

#define DL_CHECK_ARG(cond, ...)            \
    do {                                   \
        if (!(cond)) {                     \
            ::dl::set_error(__VA_ARGS__);  \
            return DL_E_ARG;               \
        }                                  \
    } while (0)

#define DL_CHECK_HIP(expr)                                                        \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::dl::set_error("%s failed: %s", #expr, hipGetErrorString(e_));       \
            return DL_E_HIP;                                                      \
        }                                                                         \
    } while (0)

#define DL_CHECK_LAUNCH() DL_CHECK_HIP(hipGetLastError())

// Raise a kernel's dynamic-LDS limit to at least `bytes`, calling the
// runtime only when the limit grows: after the first launch of a shape no
// attribute call is made, so launches issued under stream capture (HIP
// graphs) are pure launches.
inline hipError_t ensure_dynamic_lds(const void *kernel, size_t bytes) {
    static std::mutex mu;
    static std::unordered_map<const void *, size_t> done;
    if (bytes <= 65536) return hipSuccess;
    std::lock_guard<std::mutex> lock(mu);
    size_t &have = done[kernel];
    if (bytes <= have) return hipSuccess;
    const hipError_t e =
        hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) have = bytes;
    return e;
}

// ------------------------------------------------------------------ math
template <typename T> struct M;
template <> struct M<float> {
    static __device__ __forceinline__ float sqrt(float x) { return sqrtf(x); }
    static __device__ __forceinline__ float sin(float x) { return sinf(x); }
    static __device__ __forceinline__ float cos(float x) { return cosf(x); }
    static __device__ __forceinline__ float tanh(float x) { return tanhf(x); }
    static __device__ __forceinline__ float atan2(float y, float x) { return atan2f(y, x); }
    static __device__ __forceinline__ float rcbrt(float x) { return rcbrtf(x); }
};
template <> struct M<double> {
    static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
    static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
    static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
    static __device__ __forceinline__ double tanh(double x) { return ::tanh(x); }
    static __device__ __forceinline__ double atan2(double y, double x) { return ::atan2(y, x); }
    static __device__ __forceinline__ double rcbrt(double x) { return ::rcbrt(x); }
};

// Square root and reciprocal: IEEE by default; with FAST (FP32 fast path) the
// hardware v_sqrt_f32 / v_rcp_f32 instructions (about 1 ulp).
template <bool FAST, typename T>
__device__ __forceinline__ T fsqrt(T x) { return M<T>::sqrt(x); }
template <bool FAST, typename T>
__device__ __forceinline__ T frcp(T x) { return T(1) / x; }
template <>
__device__ __forceinline__ float fsqrt<true, float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <>
__device__ __forceinline__ float frcp<true, float>(float x) { return __builtin_amdgcn_rcpf(x); }

// activation value and derivative (nn.py:139,146 use sin; the quadruped MLP tanh)
template <typename T>
__device__ __forceinline__ void activate(int kind, T z, T &y, T &dy) {
    if (kind == DL_ACT_SIN) {
        y = M<T>::sin(z);
        dy = M<T>::cos(z);
    } else {
        y = M<T>::tanh(z);
        dy = T(1) - y * y;
    }
}

// -------------------------------------------------------------- reductions
constexpr int WAVE = 64;

// Lane-shifted copy of v (DPP, a VALU operand modifier: no LDS round trip).
// Lanes whose source is outside the pattern, or whose row is not in ROWS,
// read 0.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double dpp_f(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWS, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Lane K of each 16-lane row, to every lane of that row (DPP row_newbcast:K);
// the move folds into the consuming VALU op as its DPP operand.
template <int K>
__device__ __forceinline__ float row_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x150 + K, 0xF, 0xF, true));
}

// Value of lane q (wave-uniform q), to every lane.
__device__ __forceinline__ float lane_read(float v, int q) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), q));
}
__device__ __forceinline__ double lane_read(double v, int q) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), q);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), q);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes, returned to every lane.  Within each 16-lane row:
// swap neighbours, swap pairs (quad permutes), then add the mirrored half
// row and the mirrored row -- every lane then holds its row's sum; rows 1
// and 3 add the last lane of rows 0 and 2 (row_bcast:15), rows 2 and 3 the
// last lane of row 1 (row_bcast:31), leaving the total in lane 63, which is
// read back as a wave-uniform value.  All DPP, no LDS traffic.  Every lane
// of the wave must be active (a disabled lane would read as 0).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    v += dpp_f<0xB1>(v);          // quad_perm [1,0,3,2]
    v += dpp_f<0x4E>(v);          // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);         // row_half_mirror
    v += dpp_f<0x140>(v);         // row_mirror
    v += dpp_f<0x142, 0xA>(v);    // row_bcast:15 into rows 1, 3
    v += dpp_f<0x143, 0xC>(v);    // row_bcast:31 into rows 2, 3
    return lane_read(v, 63);
}

// Block-wide sum of NV values per thread; every thread gets the totals.
// `scratch` needs NV * (blockDim.x / 64) elements of shared memory.
template <typename T, int NV>
__device__ __forceinline__ void block_sum(T (&v)[NV], T *scratch) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) scratch[q * nw + wid] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        T s = T(0);
        for (int w = 0; w < nw; ++w) s += scratch[q * nw + w];  // fixed order
        v[q] = s;
    }
    __syncthreads();
}

// ------------------------------------------------------------- physics
template <typename T> struct Phys {
    T dt, g_dt, ground, drag_s, dashpot, friction;
};

template <typename T>
__host__ Phys<T> make_phys(const dl_ms_physics &p) {
    Phys<T> q;
    q.dt = T(p.dt);
    q.g_dt = T(p.dt * p.gravity);
    q.ground = T(p.ground);
    q.drag_s = T(::exp(-p.dt * p.drag));
    q.dashpot = T(p.dashpot);
    q.friction = T(p.friction);
    return q;
}

// Spring + dashpot impulse of one spring (solver_mass_spring.py:89-103):
//   L = sqrt(|d|^2 + 1e-8) + 1e-4, L* = L0 (1 + coeff a)
//   J = dt k (L - L*) / L d + c ((va - vb) . d^) d^,   d = xa - xb
// The impulse is subtracted from vertex a and added to vertex b.
template <typename T, bool FAST = false>
__device__ __forceinline__ void spring_impulse(const T d[3], const T w[3], T L0, T k, T coeff,
                                               T act, const Phys<T> &P, T J[3]) {
    const T q = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const T len = fsqrt<FAST>(q + T(1e-8)) + T(1e-4);
    const T target = L0 * (T(1) + coeff * act);
    const T A = FAST ? P.dt * (len - target) * k * frcp<FAST>(len) : P.dt * (len - target) * k / len;
    const T inv = frcp<FAST>(fsqrt<FAST>(q));
    const T dh[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
    const T vrel = w[0] * dh[0] + w[1] * dh[1] + w[2] * dh[2];
    const T C = P.dashpot * vrel;
#pragma unroll
    for (int c = 0; c < 3; ++c) J[c] = A * d[c] + C * dh[c];
}

// Adjoint of spring_impulse: given Jb = dL/dJ returns gd = dL/dd (add to xa,
// subtract from xb), gw = dL/dw (add to va, subtract from vb), ga = dL/dact.
template <typename T, bool FAST = false>
__device__ __forceinline__ void spring_impulse_adj(const T d[3], const T w[3], T L0, T k, T coeff,
                                                   T act, const Phys<T> &P, const T Jb[3],
                                                   T gd[3], T gw[3], T &ga) {
    const T q = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const T rq = fsqrt<FAST>(q + T(1e-8));
    const T len = rq + T(1e-4);
    const T ilen = frcp<FAST>(len);
    const T target = L0 * (T(1) + coeff * act);
    const T A = FAST ? P.dt * (len - target) * k * ilen : P.dt * (len - target) * k / len;
    const T n = fsqrt<FAST>(q);
    const T inv = frcp<FAST>(n);
    const T dh[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
    const T vrel = w[0] * dh[0] + w[1] * dh[1] + w[2] * dh[2];
    // term 1: A d
    const T Ab = Jb[0] * d[0] + Jb[1] * d[1] + Jb[2] * d[2];
    const T dkl = FAST ? P.dt * k * ilen : P.dt * k / len;
    const T Lb = FAST ? Ab * dkl * target * ilen : Ab * dkl * target / len;   // dA/dL = dt k L* / L^2
    const T Tb = -Ab * dkl;                      // dA/dL* = -dt k / L
    ga = Tb * L0 * coeff;
    const T lq = FAST ? Lb * frcp<FAST>(rq) : Lb / rq;   // dL/dd = d / sqrt(q + eps)
    // term 2: c vrel dh
    const T Jdh = Jb[0] * dh[0] + Jb[1] * dh[1] + Jb[2] * dh[2];
    const T vb = P.dashpot * Jdh;                // d/dvrel
    T dhb[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        dhb[c] = P.dashpot * vrel * Jb[c] + vb * w[c];
        gw[c] = vb * dh[c];
    }
    const T proj = dhb[0] * dh[0] + dhb[1] * dh[1] + dhb[2] * dh[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) gd[c] = A * Jb[c] + lq * d[c] + (dhb[c] - proj * dh[c]) * inv;
}

// Semi-implicit Euler with time-of-impact ground contact and per-axis friction
// clamp for one vertex (solver_mass_spring.py:119-148).
template <typename T>
__device__ __forceinline__ void integrate(const T x[3], const T v[3], const T vi[3],
                                          const Phys<T> &P, T xn[3], T vn[3]) {
    T ov[3] = {P.drag_s * v[0] + vi[0], P.drag_s * v[1] + P.g_dt + vi[1], P.drag_s * v[2] + vi[2]};
    const T ny = x[1] + P.dt * ov[1];
    T toi = T(0);
    T nv[3] = {ov[0], ov[1], ov[2]};
    if (ny < P.ground && ov[1] < T(-1e-4)) {
        toi = -(x[1] - P.ground) / ov[1];
        const T fr = P.friction * (-ov[1]);
        nv[0] = ov[0] < T(0) ? fmin(T(0), ov[0] + fr) : fmax(T(0), ov[0] - fr);
        nv[1] = T(0);
        nv[2] = ov[2] < T(0) ? fmin(T(0), ov[2] + fr) : fmax(T(0), ov[2] - fr);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        xn[c] = x[c] + toi * ov[c] + (P.dt - toi) * nv[c];
        vn[c] = nv[c];
    }
}

// Adjoint of integrate: given (xnb, vnb) = dL/d(x', v') returns xb, vb, vib.
template <typename T>
__device__ __forceinline__ void integrate_adj(const T x[3], const T v[3], const T vi[3],
                                              const Phys<T> &P, const T xnb[3], const T vnb[3],
                                              T xb[3], T vb[3], T vib[3]) {
    T ov[3] = {P.drag_s * v[0] + vi[0], P.drag_s * v[1] + P.g_dt + vi[1], P.drag_s * v[2] + vi[2]};
    const T ny = x[1] + P.dt * ov[1];
    T ovb[3];
    xb[0] = xnb[0]; xb[1] = xnb[1]; xb[2] = xnb[2];
    if (ny < P.ground && ov[1] < T(-1e-4)) {
        const T toi = -(x[1] - P.ground) / ov[1];
        const T fr = P.friction * (-ov[1]);
        T nv[3];
        T sx, sz;  // d nv / d ov (same-axis) and d nv / d fr sign
        T fx, fz;
        if (ov[0] < T(0)) {
            const T c = ov[0] + fr;
            nv[0] = fmin(T(0), c); sx = c < T(0) ? T(1) : T(0); fx = sx;
        } else {
            const T c = ov[0] - fr;
            nv[0] = fmax(T(0), c); sx = c > T(0) ? T(1) : T(0); fx = -sx;
        }
        if (ov[2] < T(0)) {
            const T c = ov[2] + fr;
            nv[2] = fmin(T(0), c); sz = c < T(0) ? T(1) : T(0); fz = sz;
        } else {
            const T c = ov[2] - fr;
            nv[2] = fmax(T(0), c); sz = c > T(0) ? T(1) : T(0); fz = -sz;
        }
        nv[1] = T(0);
        // x' = x + toi ov + (dt - toi) nv ; v' = nv
        T nvb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) nvb[c] = vnb[c] + (P.dt - toi) * xnb[c];
        const T toib = xnb[0] * (ov[0] - nv[0]) + xnb[1] * (ov[1] - nv[1]) + xnb[2] * (ov[2] - nv[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) ovb[c] = toi * xnb[c];
        ovb[0] += nvb[0] * sx;
        ovb[2] += nvb[2] * sz;
        const T frb = nvb[0] * fx + nvb[2] * fz;
        ovb[1] += -P.friction * frb;
        // toi = -(x.y - ground) / ov.y
        xb[1] += -toib / ov[1];
        ovb[1] += toib * (x[1] - P.ground) / (ov[1] * ov[1]);
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c) ovb[c] = vnb[c] + P.dt * xnb[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        vb[c] = P.drag_s * ovb[c];
        vib[c] = ovb[c];
    }
}

// ------------------------------------------------------------- sensors
struct Layout {
    int n_sin, run_period, vcom, vert_pad, n_tv, tv0, tv1, tv2, dup_v, dup_h, dup_c;
    double offset_scale, max_speed;
    __host__ __device__ int vstride() const { return 3 + vert_pad; }
    __host__ __device__ int n_input(int N) const {
        return n_sin + 3 * vcom + N * (3 + vert_pad) + dup_v * n_tv + dup_h + dup_c;
    }
    __host__ __device__ int vert_base() const { return n_sin + 3 * vcom; }
};

inline Layout make_layout(const dl_sensor_layout &l) {
    Layout q;
    q.n_sin = l.n_sin; q.run_period = l.run_period; q.vcom = l.vcom; q.vert_pad = l.vert_pad;
    q.n_tv = l.n_tv; q.tv0 = l.tv_comp[0]; q.tv1 = l.tv_comp[1]; q.tv2 = l.tv_comp[2];
    q.dup_v = l.dup_v; q.dup_h = l.dup_h; q.dup_c = l.dup_c;
    q.offset_scale = l.offset_scale; q.max_speed = l.max_speed;
    return q;
}

// Value of sensor slot q (taichi_env.py:150-183).  xs/vs are SoA [3][N]
// (stride N); c/vc the COM position/velocity; tgt = {tv x,y,z, th, tc}.
// The sinusoid phase is formed in double precision for every real type.
template <typename T>
__device__ __forceinline__ T sensor_value(const Layout &L, int q, int t, int offset, int N,
                                          const T *xs, const T *c, const T *vc, const T *tgt) {
    if (q < L.n_sin) {
        const double ph = 2.0 * M_PI / L.run_period * (double)(t + offset) +
                          2.0 * M_PI / L.n_sin * (double)q;
        return T(::sin(ph));
    }
    q -= L.n_sin;
    if (L.vcom) {
        if (q < 3) return q == 0 ? vc[0] : (q == 1 ? vc[1] : vc[2]);   // no dynamic private index
        q -= 3;
    }
    const int vs = L.vstride();
    if (q < N * vs) {
        const int i = q / vs, d = q - i * vs;
        if (d >= 3) return T(0);
        const T cd = d == 0 ? c[0] : (d == 1 ? c[1] : c[2]);
        return (xs[d * N + i] - cd) / T(L.offset_scale);
    }
    q -= N * vs;
    if (q < L.dup_v * L.n_tv) {
        const int m = q % L.n_tv;
        const int comp = m == 0 ? L.tv0 : (m == 1 ? L.tv1 : L.tv2);
        return tgt[comp] / T(L.max_speed);
    }
    q -= L.dup_v * L.n_tv;
    if (q < L.dup_h) return tgt[3];
    return tgt[4];
}

}  // namespace dl
