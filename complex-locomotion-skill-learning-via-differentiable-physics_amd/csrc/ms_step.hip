// Per-step mass-spring, sensor, MLP and loss kernels (the reference-shaped,
// one-launch-per-kernel-per-step API) and their adjoints.
//
// Every kernel here is deterministic: reductions run in a fixed order and
// vertex impulses are gathered through the vertex->spring CSR instead of
// being scattered with atomics.

#include "dl_common.hpp"

namespace dl {

template <typename T>
struct Topo {
    int N, S, O;
    const int32_t *sa, *sb, *sact;
    const T *L0, *k, *coeff;
    const int32_t *vptr, *vidx, *aptr, *aidx;
};

template <typename T>
static Topo<T> make_topo(const dl_ms_topology &t) {
    Topo<T> q;
    q.N = t.n_vert; q.S = t.n_spring; q.O = t.n_act;
    q.sa = t.spring_a; q.sb = t.spring_b; q.sact = t.spring_act;
    q.L0 = (const T *)t.spring_L0; q.k = (const T *)t.spring_k; q.coeff = (const T *)t.spring_coeff;
    q.vptr = t.vert_ptr; q.vidx = t.vert_idx; q.aptr = t.act_ptr; q.aidx = t.act_idx;
    return q;
}

// ------------------------------------------------------------ center / vcom
template <typename T>
__global__ void k_center(int N, const T *__restrict__ x, const T *__restrict__ v, T *center,
                         T *vcom) {
    __shared__ T red[6 * 16];
    const int k = blockIdx.x;
    const T *xk = x + (size_t)k * 3 * N;
    const T *vk = v ? v + (size_t)k * 3 * N : nullptr;
    T s[6] = {0, 0, 0, 0, 0, 0};
    const T invN = T(1) / T(N);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s[d] += xk[d * N + i] * invN;
            if (vk) s[3 + d] += vk[d * N + i] * invN;
        }
    }
    block_sum<T, 6>(s, red);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) center[k * 3 + d] = s[d];
        if (vcom)
            for (int d = 0; d < 3; ++d) vcom[k * 3 + d] = s[3 + d];
    }
}

template <typename T>
__global__ void k_center_grad(int B, int N, const T *cg, const T *vg, T *xg, T *vgout) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * N) return;
    const int k = id / N, i = id - k * N;
    const T invN = T(1) / T(N);
    for (int d = 0; d < 3; ++d) {
        if (cg) xg[(size_t)k * 3 * N + d * N + i] += cg[k * 3 + d] * invN;
        if (vg && vgout) vgout[(size_t)k * 3 * N + d * N + i] += vg[k * 3 + d] * invN;
    }
}

// ------------------------------------------------------------- height
// one thread per robot, serial over vertices like the reference (:158-173)
template <typename T>
__global__ void k_height(int B, int N, int reset, const T *x, const T *hprev, T *h, T *up) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T *y = x + (size_t)k * 3 * N + N;
    T lo = T(10), hi = T(-10);
    for (int i = 0; i < N; ++i) {
        lo = fmin(lo, y[i]);
        hi = fmax(hi, y[i]);
    }
    h[k] = reset ? lo : fmax(hprev[k], lo);
    up[k] = hi;
}

// Subgradient: the selected vertex (first minimum / first maximum) or the
// previous height receives the gradient.
template <typename T>
__global__ void k_height_grad(int B, int N, int reset, const T *x, const T *hprev, const T *hg,
                              const T *ug, T *xg, T *hpg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T *y = x + (size_t)k * 3 * N + N;
    T lo = T(10), hi = T(-10);
    int ilo = -1, ihi = -1;
    for (int i = 0; i < N; ++i) {
        if (y[i] < lo) { lo = y[i]; ilo = i; }
        if (y[i] > hi) { hi = y[i]; ihi = i; }
    }
    T *yg = xg + (size_t)k * 3 * N + N;
    if (hg) {
        if (!reset && hprev[k] >= lo) {
            if (hpg) hpg[k] += hg[k];
        } else if (ilo >= 0) {
            yg[ilo] += hg[k];
        }
    }
    if (ug && ihi >= 0) yg[ihi] += ug[k];
}

// ------------------------------------------------------------- rotation
template <typename T>
__device__ void rotation_dir(int N, int robot_id, const T *xk, const T *x0k, const T *c0k,
                             T dir[3], T &nh, T &nt) {
    if (robot_id == 100) {
        for (int d = 0; d < 3; ++d)
            dir[d] = xk[d * N + 45] + xk[d * N + 46] - xk[d * N + 34] - xk[d * N + 36];
        nh = nt = T(0);
        return;
    }
    T hc[3] = {0, 0, 0}, tc[3] = {0, 0, 0};
    nh = nt = T(0);
    for (int i = 0; i < N; ++i) {
        if (x0k[i] < c0k[0]) {
            for (int d = 0; d < 3; ++d) hc[d] += xk[d * N + i];
            nh += T(1);
        } else {
            for (int d = 0; d < 3; ++d) tc[d] += xk[d * N + i];
            nt += T(1);
        }
    }
    for (int d = 0; d < 3; ++d) dir[d] = -hc[d] * nt + tc[d] * nh;
}

template <typename T>
__global__ void k_rotation(int B, int N, int robot_id, const T *x, const T *x0, const T *c0,
                           T *rot) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    T dir[3], nh, nt;
    rotation_dir<T>(N, robot_id, x + (size_t)k * 3 * N, x0 + (size_t)k * 3 * N, c0 + k * 3, dir,
                    nh, nt);
    rot[k] = M<T>::atan2(dir[2], dir[0]);
}

template <typename T>
__global__ void k_rotation_grad(int B, int N, int robot_id, const T *x, const T *x0, const T *c0,
                                const T *rg, T *xg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T *xk = x + (size_t)k * 3 * N;
    T dir[3], nh, nt;
    rotation_dir<T>(N, robot_id, xk, x0 + (size_t)k * 3 * N, c0 + k * 3, dir, nh, nt);
    const T r2 = dir[0] * dir[0] + dir[2] * dir[2];
    if (r2 == T(0)) return;
    // d atan2(z, x) = (x dz - z dx) / (x^2 + z^2)
    const T gx = -dir[2] / r2 * rg[k], gz = dir[0] / r2 * rg[k];
    T *xgk = xg + (size_t)k * 3 * N;
    if (robot_id == 100) {
        const int ids[4] = {45, 46, 34, 36};
        const T sg[4] = {1, 1, -1, -1};
        for (int q = 0; q < 4; ++q) {
            xgk[ids[q]] += sg[q] * gx;
            xgk[2 * N + ids[q]] += sg[q] * gz;
        }
        return;
    }
    const T *x0k = x0 + (size_t)k * 3 * N;
    for (int i = 0; i < N; ++i) {
        const T w = x0k[i] < c0[k * 3] ? -nt : nh;
        xgk[i] += w * gx;
        xgk[2 * N + i] += w * gz;
    }
}

// ------------------------------------------------------------- sensors
template <typename T>
__global__ void k_nn_input(Layout L, int N, int I, int t, int offset, const T *x, const T *v,
                           const T *center, const T *tv, const T *th, const T *tc, T *in) {
    const int k = blockIdx.y;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *vc = reinterpret_cast<T *>(smem_raw);
    const T *xk = x + (size_t)k * 3 * N;
    if (L.vcom) {
        // COM velocity (fixed-order block sum)
        T s[3] = {0, 0, 0};
        const T invN = T(1) / T(N);
        for (int i = threadIdx.x; i < N; i += blockDim.x)
            for (int d = 0; d < 3; ++d) s[d] += v[(size_t)k * 3 * N + d * N + i] * invN;
        block_sum<T, 3>(s, vc + 3);
        if (threadIdx.x == 0)
            for (int d = 0; d < 3; ++d) vc[d] = s[d];
        __syncthreads();
    }
    T tgt[5] = {tv[k * 3 + 0], tv[k * 3 + 1], tv[k * 3 + 2], th[k], tc ? tc[k] : T(0)};
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < I) in[(size_t)k * I + q] = sensor_value<T>(L, q, t, offset, N, xk, center + k * 3, vc, tgt);
}

template <typename T>
__global__ void k_nn_input_grad(Layout L, int N, int I, const T *ing, T *xg, T *vg, T *cg) {
    // one block per robot: vertex slots -> x, center; vcom slots -> v
    const int k = blockIdx.x;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    const T *g = ing + (size_t)k * I;
    const int base = L.vert_base(), vs = L.vstride();
    const T inv = T(1) / T(L.offset_scale);
    T s[3] = {0, 0, 0};
    const T invN = T(1) / T(N);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            const T gi = g[base + i * vs + d] * inv;
            if (xg) xg[(size_t)k * 3 * N + d * N + i] += gi;
            s[d] += gi;
            if (L.vcom && vg) vg[(size_t)k * 3 * N + d * N + i] += g[L.n_sin + d] * invN;
        }
    }
    block_sum<T, 3>(s, red);
    if (threadIdx.x == 0 && cg)
        for (int d = 0; d < 3; ++d) cg[k * 3 + d] -= s[d];
}

// ------------------------------------------------------------- MLP
// Block per robot.  hidden j handled by thread (j, chunk); W1 rows read
// coalesced across j through the transposed copy computed on the fly from
// W1 [H][I] (row-major, reference layout): thread j reads W1[j][i], which is
// strided, so the per-step API stages W1 tiles through LDS.
template <typename T>
__global__ void k_mlp_fwd(int act, int I, int H, int O, const T *in, const T *W1, const T *b1,
                          const T *W2, const T *b2, T *hid, T *out, T *z1save, T *z2save) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *xin = reinterpret_cast<T *>(smem_raw);   // [I]
    T *h = xin + I;                             // [H]
    const int k = blockIdx.x;
    for (int i = threadIdx.x; i < I; i += blockDim.x) xin[i] = in[(size_t)k * I + i];
    __syncthreads();
    // layer 1: one wave per hidden unit at a time, lanes over inputs (coalesced W1 rows)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = wid; j < H; j += nw) {
        T acc = T(0);
        for (int i = lane; i < I; i += 64) acc += W1[(size_t)j * I + i] * xin[i];
        acc = wave_sum(acc);
        if (lane == 0) {
            T y, dy;
            const T z = acc + b1[j];
            activate<T>(act, z, y, dy);
            h[j] = y;
            if (hid) hid[(size_t)k * H + j] = y;
            if (z1save) z1save[(size_t)k * H + j] = z;
        }
    }
    __syncthreads();
    for (int o = wid; o < O; o += nw) {
        T acc = T(0);
        for (int j = lane; j < H; j += 64) acc += W2[(size_t)o * H + j] * h[j];
        acc = wave_sum(acc);
        if (lane == 0) {
            T y, dy;
            const T z = acc + b2[o];
            activate<T>(act, z, y, dy);
            out[(size_t)k * O + o] = y;
            if (z2save) z2save[(size_t)k * O + o] = z;
        }
    }
}

// Adjoint, stage 1 (block per robot): recompute, then z2b = g*act'(z2),
// h_b = W2^T z2b, z1b = h_b*act'(z1), in_b = W1^T z1b.  Writes z1b, z2b, h
// into work [B][2H+O] for the deterministic weight-gradient stage.
template <typename T>
__global__ void k_mlp_bwd1(int act, int I, int H, int O, const T *in, const T *W1, const T *b1,
                           const T *W2, const T *b2, const T *og, T *ing, T *work) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *xin = reinterpret_cast<T *>(smem_raw);   // [I]
    T *h = xin + I;                             // [H]
    T *dh = h + H;                              // [H]
    T *z2b = dh + H;                            // [O]
    T *z1b = z2b + O;                           // [H]
    const int k = blockIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = threadIdx.x; i < I; i += blockDim.x) xin[i] = in[(size_t)k * I + i];
    __syncthreads();
    for (int j = wid; j < H; j += nw) {
        T acc = T(0);
        for (int i = lane; i < I; i += 64) acc += W1[(size_t)j * I + i] * xin[i];
        acc = wave_sum(acc);
        if (lane == 0) {
            T y, dy;
            activate<T>(act, acc + b1[j], y, dy);
            h[j] = y;
            dh[j] = dy;
        }
    }
    __syncthreads();
    for (int o = wid; o < O; o += nw) {
        T acc = T(0);
        for (int j = lane; j < H; j += 64) acc += W2[(size_t)o * H + j] * h[j];
        acc = wave_sum(acc);
        if (lane == 0) {
            T y, dy;
            activate<T>(act, acc + b2[o], y, dy);
            z2b[o] = og[(size_t)k * O + o] * dy;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
        T acc = T(0);
        for (int o = 0; o < O; ++o) acc += W2[(size_t)o * H + j] * z2b[o];
        z1b[j] = acc * dh[j];
    }
    __syncthreads();
    if (ing) {
        for (int i = threadIdx.x; i < I; i += blockDim.x) {
            T acc = T(0);
            for (int j = 0; j < H; ++j) acc += W1[(size_t)j * I + i] * z1b[j];
            ing[(size_t)k * I + i] += acc;
        }
    }
    T *wk = work + (size_t)k * (2 * H + O);
    for (int j = threadIdx.x; j < H; j += blockDim.x) {
        wk[j] = z1b[j];
        wk[H + j] = h[j];
    }
    for (int o = threadIdx.x; o < O; o += blockDim.x) wk[2 * H + o] = z2b[o];
}

// Adjoint, stage 2: weight gradients summed over the batch in order k = 0..B-1.
template <typename T>
__global__ void k_mlp_bwd2(int B, int I, int H, int O, const T *in, const T *work, T *W1g, T *b1g,
                           T *W2g, T *b2g) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = H * I, n2 = O * H;
    const int st = 2 * H + O;
    if (id < n1) {
        const int j = id / I, i = id - j * I;
        T acc = T(0);
        for (int k = 0; k < B; ++k) acc += work[(size_t)k * st + j] * in[(size_t)k * I + i];
        W1g[id] += acc;
    } else if (id < n1 + n2) {
        const int q = id - n1, o = q / H, j = q - o * H;
        T acc = T(0);
        for (int k = 0; k < B; ++k) acc += work[(size_t)k * st + 2 * H + o] * work[(size_t)k * st + H + j];
        W2g[q] += acc;
    } else if (id < n1 + n2 + H) {
        const int j = id - n1 - n2;
        T acc = T(0);
        for (int k = 0; k < B; ++k) acc += work[(size_t)k * st + j];
        b1g[j] += acc;
    } else if (id < n1 + n2 + H + O) {
        const int o = id - n1 - n2 - H;
        T acc = T(0);
        for (int k = 0; k < B; ++k) acc += work[(size_t)k * st + 2 * H + o];
        b2g[o] += acc;
    }
}


// ------------------------------------------------------------- MLP on MFMA
// FP32 batched MLP with the two big products on the matrix cores
// (v_mfma_f32_16x16x4f32): a workgroup of 4 wavefronts takes a tile of 16
// robots; wavefront w computes hidden units 16w .. 16w+15 of layer 1
// (Z1 = X W1^T, K = I in steps of 4), and in the adjoint the input gradient
// columns (in_b = z1b W1, K = H).  The robots' input rows are staged in LDS
// (odd row pitch); W1 operands are read straight from global memory (it is
// small and stays in L2: staging all of it per workgroup cost more than the
// products).  Layer 2 (O <= 16 outputs) runs on VALU from the LDS hidden
// tile, four lanes per output.
//
// MFMA 16x16x4 (wave64) operand layout: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; result register j of lane l is D[4 (l / 16) + j][l % 16].
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int MT = 16;            // robots per workgroup tile
constexpr int MFMA_HMAX = 64, MFMA_IMAX = 256, MFMA_OMAX = 16;

__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// LDS: x tile [MT][IP] (IP = I rounded up to a multiple of 4, plus one: odd
// pitch; the columns past I are zero), then the hidden tile and friends.
__device__ __forceinline__ int mlp_pitch(int I) { return ((I + 3) & ~3) + 1; }

// Wavefront w stages rows w, w + 4, .. of the tile (blockDim 256), all its
// loads issued before the first LDS store.
__device__ __forceinline__ void mlp_stage(int B, int I, int IP, int r0, const float *in,
                                          float *xs) {
    constexpr int RW = MT / 4, CW = MFMA_IMAX / 64;
    const int I4 = (I + 3) & ~3, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float v[RW][CW];
#pragma unroll
    for (int a = 0; a < RW; ++a)
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int m = w + 4 * a, i = lane + 64 * c;
            v[a][c] = (r0 + m < B && i < I) ? in[(size_t)(r0 + m) * I + i] : 0.f;
        }
#pragma unroll
    for (int a = 0; a < RW; ++a)
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int i = lane + 64 * c;
            if (i < I4) xs[(w + 4 * a) * IP + i] = v[a][c];
        }
}

// Z1 tile of wavefront w: D[m][n] = sum_k X[m][k] W1[16w + n][k], W1 rows
// from global memory (16 w + 15 < H).  Every W1 operand of the tile is
// loaded before the first product (one round trip to L2 instead of one per
// few k steps); the k-step guards are uniform across the wavefront.
__device__ __forceinline__ v4f mlp_layer1(int I, int IP, const float *xs, const float *W1, int w,
                                          int lane) {
    constexpr int KS = MFMA_IMAX / 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    const int r = lane & 15, kq = lane >> 4;
    const float *xa = xs + r * IP + kq;
    const float *wb = W1 + (size_t)(16 * w + r) * I + kq;
    float wv[KS];
#pragma unroll
    for (int k = 0; k < KS; ++k) wv[k] = 4 * k + kq < I ? wb[4 * k] : 0.f;
#pragma unroll
    for (int k = 0; k < KS; ++k)
        if (4 * k < I) acc = mfma16x16x4(xa[4 * k], wv[k], acc);
    return acc;
}

// Layer 2 of the tile from the LDS hidden rows hs [MT][HP]: four lanes per
// (robot, output) pair, each summing a quarter of the hidden units, combined
// by lane shuffles; returns z2 (valid in every lane of the four) for the pair
// e = (thread index) / 4 + 64 round.
__device__ __forceinline__ float mlp_layer2_part(int H, int HP, int O, const float *W2,
                                                 const float *b2, const float *hs, int e) {
    const int m = e / O, o = e - m * O, q = threadIdx.x & 3, hq = H >> 2;
    float z = 0.f;
    if (m < MT)
        for (int n = q * hq; n < (q + 1) * hq; ++n) z += W2[(size_t)o * H + n] * hs[m * HP + n];
    z += __shfl_xor(z, 1);
    z += __shfl_xor(z, 2);
    return z + (m < MT ? b2[o] : 0.f);
}

__global__ void __launch_bounds__(256) k_mlp_fwd_mfma(int B, int act, int I, int H, int O,
                                                      const float *in, const float *W1,
                                                      const float *b1, const float *W2,
                                                      const float *b2, float *hid, float *out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int IP = mlp_pitch(I);
    float *xs = reinterpret_cast<float *>(smem_raw);
    float *hs = xs + MT * IP;                     // [MT][H + 1]
    const int r0 = blockIdx.x * MT, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    mlp_stage(B, I, IP, r0, in, xs);
    __syncthreads();
    if (16 * w < H) {
        const v4f z = mlp_layer1(I, IP, xs, W1, w, lane);
        const int n = 16 * w + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * (lane >> 4) + j;
            float y, dy;
            activate<float>(act, z[j] + b1[n], y, dy);
            hs[m * (H + 1) + n] = y;
            if (r0 + m < B && hid) hid[(size_t)(r0 + m) * H + n] = y;
        }
    }
    __syncthreads();
    for (int e0 = 0; e0 < MT * O; e0 += blockDim.x >> 2) {
        const int e = e0 + (threadIdx.x >> 2);
        const float z = mlp_layer2_part(H, H + 1, O, W2, b2, hs, e);
        const int m = e / O, o = e - m * O;
        float y, dy;
        activate<float>(act, z, y, dy);
        if ((threadIdx.x & 3) == 0 && m < MT && r0 + m < B) out[(size_t)(r0 + m) * O + o] = y;
    }
}

// Adjoint, stage 1 (tile of 16 robots): recompute z1 (MFMA), h, z2; z2b =
// og act'(z2), z1b = (z2b W2) act'(z1); in_b += z1b W1 (MFMA, K = H); the
// rows z1b, h, z2b go to work [B][2H+O] for the weight-gradient stage.
__global__ void __launch_bounds__(256) k_mlp_bwd1_mfma(int B, int act, int I, int H, int O,
                                                       const float *in, const float *W1,
                                                       const float *b1, const float *W2,
                                                       const float *b2, const float *og,
                                                       float *ing, float *work) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int IP = mlp_pitch(I), HP = H + 1;
    float *xs = reinterpret_cast<float *>(smem_raw);
    float *hs = xs + MT * IP;                     // h    [MT][HP]
    float *ds = hs + MT * HP;                     // act' [MT][HP], then z1b
    float *zb = ds + MT * HP;                     // z2b  [MT][O]
    const int r0 = blockIdx.x * MT, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int st = 2 * H + O;
    mlp_stage(B, I, IP, r0, in, xs);
    __syncthreads();
    if (16 * w < H) {
        const v4f z = mlp_layer1(I, IP, xs, W1, w, lane);
        const int n = 16 * w + (lane & 15);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * (lane >> 4) + j;
            float y, dy;
            activate<float>(act, z[j] + b1[n], y, dy);
            hs[m * HP + n] = y;
            ds[m * HP + n] = dy;
        }
    }
    __syncthreads();
    for (int e0 = 0; e0 < MT * O; e0 += blockDim.x >> 2) {
        const int e = e0 + (threadIdx.x >> 2);
        const float z = mlp_layer2_part(H, HP, O, W2, b2, hs, e);
        const int m = e / O, o = e - m * O;
        float y, dy;
        activate<float>(act, z, y, dy);
        if ((threadIdx.x & 3) == 0 && m < MT)
            zb[m * O + o] = r0 + m < B ? og[(size_t)(r0 + m) * O + o] * dy : 0.f;
    }
    __syncthreads();
    {
        // (robot, hidden) pairs e = threadIdx.x + 256 c; W2 operands first
        constexpr int NC = MT * MFMA_HMAX / 256;
        float w2v[NC][MFMA_OMAX];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int e = threadIdx.x + 256 * c, n = e % H;
#pragma unroll
            for (int o = 0; o < MFMA_OMAX; ++o)
                w2v[c][o] = o < O && e < MT * H ? W2[(size_t)o * H + n] : 0.f;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int e = threadIdx.x + 256 * c;
            if (e < MT * H) {
                const int m = e / H, n = e - m * H;
                float hb = 0.f;
#pragma unroll
                for (int o = 0; o < MFMA_OMAX; ++o)
                    if (o < O) hb += w2v[c][o] * zb[m * O + o];
                const float z1b = hb * ds[m * HP + n];
                ds[m * HP + n] = z1b;
                if (r0 + m < B) {
                    work[(size_t)(r0 + m) * st + n] = z1b;
                    work[(size_t)(r0 + m) * st + H + n] = hs[m * HP + n];
                }
            }
        }
    }
    for (int e = threadIdx.x; e < MT * O; e += blockDim.x) {
        const int m = e / O, o = e - m * O;
        if (r0 + m < B) work[(size_t)(r0 + m) * st + 2 * H + o] = zb[m * O + o];
    }
    __syncthreads();
    if (!ing) return;
    // in_b[m][i] = sum_n z1b[m][n] W1[n][i]: A = z1b (16 x H), B = W1 (H x I)
    // column tiles nt = w, w + 4, .. (at most MFMA_IMAX / 64 per wavefront);
    // the W1 operands and the in_b values to add to of every tile are loaded
    // before the first product
    const int r = lane & 15, kq = lane >> 4;
    constexpr int KS = MFMA_HMAX / 4, NT = MFMA_IMAX / 64;
    float wv[NT][KS], gv[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int i = (w + 4 * t) * 16 + r;
        const bool iok = i < I;
        const float *wc = W1 + (size_t)kq * I + (iok ? i : 0);
#pragma unroll
        for (int k = 0; k < KS; ++k) wv[t][k] = 4 * k < H && iok ? wc[(size_t)4 * k * I] : 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * kq + j;
            gv[t][j] = r0 + m < B && iok ? ing[(size_t)(r0 + m) * I + i] : 0.f;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int i = (w + 4 * t) * 16 + r;
        if ((w + 4 * t) * 16 >= I) break;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < KS; ++k)
            if (4 * k < H) acc = mfma16x16x4(ds[r * HP + 4 * k + kq], wv[t][k], acc);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int m = 4 * kq + j;
            if (r0 + m < B && i < I) ing[(size_t)(r0 + m) * I + i] = gv[t][j] + acc[j];
        }
    }
}

// Adjoint, stage 2: every weight gradient as an MFMA product over the batch
// (K = B), one 16 x 16 output tile per workgroup: W1g = z1b^T X (H/16 x
// ceil(I/16) tiles), W2g = z2b^T h (H/16 tiles, rows o < O), b1g = 1^T z1b
// (H/16 tiles, row 0) and b2g = 1^T z2b (one tile, row 0).  The batch is
// split over the 4 wavefronts in contiguous quarters and the partial tiles
// are added in wavefront order, so the sums are deterministic.  work rows
// are [z1b (H) | h (H) | z2b (O)].
__global__ void __launch_bounds__(256) k_mlp_bwd2_mfma(int B, int I, int H, int O, int ntI,
                                                       const float *in, const float *work,
                                                       float *W1g, float *b1g, float *W2g,
                                                       float *b2g) {
    __shared__ float part[4][256];
    const int st = 2 * H + O, nh = H / 16;
    int t = blockIdx.x, kind, col0;
    if (t < nh * ntI) {
        kind = 0;
    } else if ((t -= nh * ntI) < nh) {
        kind = 1;
    } else if ((t -= nh) < nh) {
        kind = 2;
    } else {
        kind = 3;
        t = 0;
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    const int q = (((B + 3) >> 2) + 3) & ~3;      // batch rows per wavefront, multiple of 4
    const int k_lo = w * q, k_hi = min(B, k_lo + q);
    // A[m][b] and B[b][n] of this tile for lane row / column r
    int a_off = -1, b_off = -1, b_lim = 0;
    float a_one = 0.f;
    if (kind == 0) {                 // A = z1b[b][16 mt + m], B = X[b][16 nt + n]
        const int mt = t / ntI, nt = t - mt * ntI;
        a_off = mt * 16 + r;
        col0 = nt * 16;
        b_lim = I;
    } else if (kind == 1) {          // A = z2b[b][m] (m < O), B = h[b][16 t + n]
        a_off = r < O ? 2 * H + r : -1;
        col0 = t * 16;
        b_lim = H;
    } else if (kind == 2) {          // A = 1 (row 0), B = z1b[b][16 t + n]
        a_one = r == 0 ? 1.f : 0.f;
        col0 = t * 16;
        b_lim = H;
    } else {                         // A = 1 (row 0), B = z2b[b][n] (n < O)
        a_one = r == 0 ? 1.f : 0.f;
        col0 = 0;
        b_lim = O;
    }
    const int n = col0 + r;
    const bool n_ok = n < b_lim;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int k0 = k_lo; k0 < k_hi; k0 += 4) {
        const int b = k0 + kq;
        const bool ok = b < k_hi;
        const float *row = work + (size_t)b * st;
        float a = a_one, x = 0.f;
        if (ok && a_off >= 0) a = row[a_off];
        if (!ok) a = 0.f;
        if (ok && n_ok) {
            if (kind == 0) x = in[(size_t)b * I + n];
            else if (kind == 1) x = row[H + n];
            else if (kind == 2) x = row[n];
            else x = row[2 * H + n];
        }
        acc = mfma16x16x4(a, x, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) part[w][(4 * kq + j) * 16 + r] = acc[j];
    __syncthreads();
    const int m = threadIdx.x >> 4, c = threadIdx.x & 15;
    const float v = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                    part[3][threadIdx.x];
    const int cc = col0 + c;
    if (kind == 0) {
        const int mt = t / ntI;
        if (cc < I) W1g[(size_t)(mt * 16 + m) * I + cc] += v;
    } else if (kind == 1) {
        if (m < O) W2g[(size_t)m * H + cc] += v;
    } else if (kind == 2) {
        if (m == 0) b1g[cc] += v;
    } else if (m == 0 && cc < O) {
        b2g[cc] += v;
    }
}

static inline bool mlp_mfma_ok(int I, int H, int O) {
    return H % 16 == 0 && H <= MFMA_HMAX && I <= MFMA_IMAX && O <= MFMA_OMAX;
}

// DL_OPT_MLP_MFMA: 0 VALU kernels only; 1 the faster kernel of each stage by
// measurement at 1024 x 179 x 64 x 4 (MFMA forward and weight-gradient stage,
// VALU input-adjoint stage: one workgroup per robot keeps more of the chip
// busy than the 16-robot MFMA tiles there); 2 every stage on MFMA.
static int g_mlp_mfma = 1;

// ------------------------------------------------------------- advance
template <typename T>
__device__ __forceinline__ void load3(const T *a, int N, int i, T r[3]) {
    r[0] = a[i]; r[1] = a[N + i]; r[2] = a[2 * N + i];
}

template <typename T>
__device__ __forceinline__ void spring_of(const Topo<T> &tp, const T *xk, const T *vk,
                                          const T *ak, const Phys<T> &P, int s, T J[3]) {
    const int a = tp.sa[s], b = tp.sb[s];
    T xa[3], xb[3], va[3], vb[3];
    load3(xk, tp.N, a, xa); load3(xk, tp.N, b, xb);
    load3(vk, tp.N, a, va); load3(vk, tp.N, b, vb);
    const T d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
    const T w[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
    spring_impulse<T>(d, w, tp.L0[s], tp.k[s], tp.coeff[s], ak[tp.sact[s]], P, J);
}

template <typename T>
__device__ __forceinline__ void gather_vinc(const Topo<T> &tp, const T *xk, const T *vk,
                                            const T *ak, const Phys<T> &P, int i, T vi[3]) {
    vi[0] = vi[1] = vi[2] = T(0);
    for (int e = tp.vptr[i]; e < tp.vptr[i + 1]; ++e) {
        const int ent = tp.vidx[e];
        T J[3];
        spring_of<T>(tp, xk, vk, ak, P, ent >> 1, J);
        const T sg = (ent & 1) ? T(1) : T(-1);
        vi[0] += sg * J[0]; vi[1] += sg * J[1]; vi[2] += sg * J[2];
    }
}

template <typename T>
__global__ void k_advance(Topo<T> tp, Phys<T> P, int B, const T *x, const T *v, const T *act,
                          T *xo, T *vo) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * tp.N) return;
    const int k = id / tp.N, i = id - k * tp.N;
    const size_t off = (size_t)k * 3 * tp.N;
    const T *xk = x + off, *vk = v + off, *ak = act + (size_t)k * tp.O;
    T vi[3], xi[3], vv[3], xn[3], vn[3];
    gather_vinc<T>(tp, xk, vk, ak, P, i, vi);
    load3(xk, tp.N, i, xi);
    load3(vk, tp.N, i, vv);
    integrate<T>(xi, vv, vi, P, xn, vn);
    for (int d = 0; d < 3; ++d) {
        xo[off + d * tp.N + i] = xn[d];
        vo[off + d * tp.N + i] = vn[d];
    }
}

// Adjoint of k_advance, block per robot.  LDS: vib [3N], gs [7S]
template <typename T>
__global__ void k_advance_grad(Topo<T> tp, Phys<T> P, const T *x, const T *v, const T *act,
                               const T *xg1, const T *vg1, T *xg, T *vg, T *ag) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int N = tp.N, S = tp.S;
    T *vib = reinterpret_cast<T *>(smem_raw);  // [3][N]
    T *gs = vib + 3 * N;                       // [7][S]: gd, gw, ga
    const int k = blockIdx.x;
    const size_t off = (size_t)k * 3 * N;
    const T *xk = x + off, *vk = v + off, *ak = act + (size_t)k * tp.O;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        T vi[3], xi[3], vv[3], xnb[3], vnb[3], xb[3], vb[3], vb2[3];
        gather_vinc<T>(tp, xk, vk, ak, P, i, vi);
        load3(xk, N, i, xi);
        load3(vk, N, i, vv);
        load3(xg1 + off, N, i, xnb);
        load3(vg1 + off, N, i, vnb);
        integrate_adj<T>(xi, vv, vi, P, xnb, vnb, xb, vb, vb2);
        for (int d = 0; d < 3; ++d) {
            xg[off + d * N + i] += xb[d];
            vg[off + d * N + i] += vb[d];
            vib[d * N + i] = vb2[d];
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
        const int a = tp.sa[s], b = tp.sb[s];
        T xa[3], xb[3], va[3], vb[3];
        load3(xk, N, a, xa); load3(xk, N, b, xb);
        load3(vk, N, a, va); load3(vk, N, b, vb);
        const T d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        const T w[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
        const T Jb[3] = {vib[b] - vib[a], vib[N + b] - vib[N + a], vib[2 * N + b] - vib[2 * N + a]};
        T gd[3], gw[3], ga;
        spring_impulse_adj<T>(d, w, tp.L0[s], tp.k[s], tp.coeff[s], ak[tp.sact[s]], P, Jb, gd, gw, ga);
        for (int c = 0; c < 3; ++c) {
            gs[c * S + s] = gd[c];
            gs[(3 + c) * S + s] = gw[c];
        }
        gs[6 * S + s] = ga;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        T sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
        for (int e = tp.vptr[i]; e < tp.vptr[i + 1]; ++e) {
            const int ent = tp.vidx[e], s = ent >> 1;
            const T sg = (ent & 1) ? T(-1) : T(1);
            for (int c = 0; c < 3; ++c) {
                sx[c] += sg * gs[c * S + s];
                sv[c] += sg * gs[(3 + c) * S + s];
            }
        }
        for (int d = 0; d < 3; ++d) {
            xg[off + d * N + i] += sx[d];
            vg[off + d * N + i] += sv[d];
        }
    }
    if (ag) {
        for (int o = threadIdx.x; o < tp.O; o += blockDim.x) {
            T s = T(0);
            for (int e = tp.aptr[o]; e < tp.aptr[o + 1]; ++e) s += gs[6 * S + tp.aidx[e]];
            ag[(size_t)k * tp.O + o] += s;
        }
    }
}

// ------------------------------------------------------------- losses
// compute_loss_velocity (taichi_env.py:185-211); single block, robot per
// thread, fixed order.  Entries t >= Tw read as zero.
template <typename T>
__device__ __forceinline__ T cat(const T *c, int Tw, int B, int t, int k, int d) {
    return t < Tw ? c[((size_t)t * B + k) * 3 + d] : T(0);
}

template <typename T>
__global__ void k_loss_velocity(int B, int Tw, int steps, int rp, int tp, const T *c, const T *tv,
                                T *loss, T *loss_b) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    T tot[1] = {T(0)};
    for (int k = threadIdx.x; k < B; k += blockDim.x) {
        T s = T(0);
        for (int t = rp; t <= steps; ++t) {
            if (t % tp > rp) {
                const T lx = cat(c, Tw, B, t, k, 0) - cat(c, Tw, B, t - rp, k, 0) -
                             tv[((size_t)(t - rp) * B + k) * 3 + 0];
                const T lz = cat(c, Tw, B, t, k, 2) - cat(c, Tw, B, t - rp, k, 2) -
                             tv[((size_t)(t - rp) * B + k) * 3 + 2];
                s += (lx * lx + lz * lz) / T(B) / T(steps) * T(100);
            }
        }
        if (loss_b) loss_b[k] = s * T(B);
        tot[0] += s;
    }
    block_sum<T, 1>(tot, red);
    if (threadIdx.x == 0) loss[0] = tot[0];
}

template <typename T>
__global__ void k_loss_velocity_grad(int B, int Tw, int steps, int rp, int tp, const T *c,
                                     const T *tv, T g, T *cg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    for (int t = rp; t <= steps; ++t) {
        if (t % tp > rp) {
            for (int d = 0; d < 3; d += 2) {
                const T l = cat(c, Tw, B, t, k, d) - cat(c, Tw, B, t - rp, k, d) -
                            tv[((size_t)(t - rp) * B + k) * 3 + d];
                const T gl = g * T(2) * l / T(B) / T(steps) * T(100);
                if (t < Tw) cg[((size_t)t * B + k) * 3 + d] += gl;
                if (t - rp < Tw) cg[((size_t)(t - rp) * B + k) * 3 + d] -= gl;
            }
        }
    }
}

// compute_loss_rotation, active loop (taichi_env.py:244-252); block per robot.
template <typename T>
__global__ void k_loss_rotation(int B, int N, int Tw, int Tx, int steps, int rp, int tp,
                                const T *x, const T *c, const T *th, T *lossb_tmp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    const int k = blockIdx.x;
    T s[1] = {T(0)};
    for (int t = 1; t <= steps; ++t) {
        if (t % tp <= rp) continue;
        const int t1 = t - rp;
        const T ang = th[(size_t)t1 * B + k] * T(72.0) / T(180.0) * T(3.1415926535);
        const T ca = M<T>::cos(ang), sa = M<T>::sin(ang);
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            T p0[3], p1[3];
            for (int d = 0; d < 3; ++d) {
                p0[d] = (t < Tx ? x[((size_t)t * B + k) * 3 * N + d * N + i] : T(0)) - cat(c, Tw, B, t, k, d);
                p1[d] = (t1 < Tx ? x[((size_t)t1 * B + k) * 3 * N + d * N + i] : T(0)) - cat(c, Tw, B, t1, k, d);
            }
            const T q0 = p0[0] - (ca * p1[0] - sa * p1[2]);
            const T q1 = p0[1] - p1[1];
            const T q2 = p0[2] - (sa * p1[0] + ca * p1[2]);
            s[0] += (q0 * q0 + q1 * q1 + q2 * q2) / T(steps) / T(40);
        }
    }
    block_sum<T, 1>(s, red);
    if (threadIdx.x == 0) lossb_tmp[k] = s[0];
}

template <typename T>
__global__ void k_loss_sum(int B, const T *part, T *loss, T *loss_b) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    T s[1] = {T(0)};
    for (int k = threadIdx.x; k < B; k += blockDim.x) {
        s[0] += part[k];
        if (loss_b) loss_b[k] = part[k] * T(B);
    }
    block_sum<T, 1>(s, red);
    if (threadIdx.x == 0) loss[0] = s[0];
}

template <typename T>
__global__ void k_loss_rotation_grad(int B, int N, int Tw, int Tx, int steps, int rp, int tp,
                                     const T *x, const T *c, const T *th, T g, T *xg, T *cg) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    const int k = blockIdx.x;
    for (int t = 1; t <= steps; ++t) {
        if (t % tp <= rp) continue;
        const int t1 = t - rp;
        const T ang = th[(size_t)t1 * B + k] * T(72.0) / T(180.0) * T(3.1415926535);
        const T ca = M<T>::cos(ang), sa = M<T>::sin(ang);
        T s[6] = {0, 0, 0, 0, 0, 0};
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            T p0[3], p1[3];
            for (int d = 0; d < 3; ++d) {
                p0[d] = (t < Tx ? x[((size_t)t * B + k) * 3 * N + d * N + i] : T(0)) - cat(c, Tw, B, t, k, d);
                p1[d] = (t1 < Tx ? x[((size_t)t1 * B + k) * 3 * N + d * N + i] : T(0)) - cat(c, Tw, B, t1, k, d);
            }
            const T sc = g * T(2) / T(steps) / T(40);
            const T q[3] = {(p0[0] - (ca * p1[0] - sa * p1[2])) * sc, (p0[1] - p1[1]) * sc,
                            (p0[2] - (sa * p1[0] + ca * p1[2])) * sc};
            // R^T q
            const T r[3] = {ca * q[0] + sa * q[2], q[1], -sa * q[0] + ca * q[2]};
            for (int d = 0; d < 3; ++d) {
                if (t < Tx) xg[((size_t)t * B + k) * 3 * N + d * N + i] += q[d];
                if (t1 < Tx) xg[((size_t)t1 * B + k) * 3 * N + d * N + i] -= r[d];
                s[d] += q[d];
                s[3 + d] += r[d];
            }
        }
        block_sum<T, 6>(s, red);
        if (threadIdx.x == 0) {
            for (int d = 0; d < 3; ++d) {
                if (t < Tw) cg[((size_t)t * B + k) * 3 + d] -= s[d];
                if (t1 < Tw) cg[((size_t)t1 * B + k) * 3 + d] += s[3 + d];
            }
        }
    }
}

// Windowed COM loss of the quadruped task: mean over robots and over states
// t in (T - W, T] of |vc_xz - tv_xz|^2 + (c_y - th)^2.  center/vcom [T+1][B][3],
// tv [B][3], th [B].  loss_b[k] = per-robot mean over the window.
template <typename T>
__global__ void k_loss_task(int B, int Tn, int W, const T *c, const T *vc, const T *tv,
                            const T *th, T *loss, T *loss_b) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    T tot[1] = {T(0)};
    for (int k = threadIdx.x; k < B; k += blockDim.x) {
        T s = T(0);
        for (int t = Tn - W + 1; t <= Tn; ++t) {
            const size_t o = ((size_t)t * B + k) * 3;
            const T a = vc[o] - tv[k * 3], b = vc[o + 2] - tv[k * 3 + 2], h = c[o + 1] - th[k];
            s += a * a + b * b + h * h;
        }
        s /= T(W);
        if (loss_b) loss_b[k] = s;
        tot[0] += s;
    }
    block_sum<T, 1>(tot, red);
    if (threadIdx.x == 0) loss[0] = tot[0] / T(B);
}

template <typename T>
__global__ void k_loss_task_grad(int B, int Tn, int W, const T *c, const T *vc, const T *tv,
                                 const T *th, T g, T *cg, T *vcg) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * W) return;
    const int k = id % B, t = Tn - W + 1 + id / B;
    const size_t o = ((size_t)t * B + k) * 3;
    const T sc = g * T(2) / (T(B) * T(W));
    vcg[o] += sc * (vc[o] - tv[k * 3]);
    vcg[o + 2] += sc * (vc[o + 2] - tv[k * 3 + 2]);
    cg[o + 1] += sc * (c[o + 1] - th[k]);
}


// ------------------------------------------------ height / pose / actuation / crawl
// Rows t >= Tw of a field the rollout did not write read as zero (as in the
// other losses).  Per-robot partial sums go to `part` [B]; k_loss_sum adds
// them in a fixed order.

// compute_loss_height (taichi_env.py:213-222): t in [1, steps] with
// t % jp == jp - 1 and th > 0.1 + 1e-6.  hmax[k]: the largest contributing
// height of robot k (-inf when none; current_max_height).
template <typename T>
__global__ void k_loss_height(int B, int Tw, int steps, int jp, const T *h, const T *th, T *part,
                              T *hmax) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T sc = T(100) / T(B) / T(steps / jp);
    T s = T(0), mx = -INFINITY;
    for (int t = jp - 1; t <= steps; t += jp) {
        if (t < 1) continue;
        const T tg = th[(size_t)t * B + k];
        if (!(tg > T(0.1) + T(1e-6))) continue;
        const T hv = t < Tw ? h[(size_t)t * B + k] : T(0);
        s += (hv - tg) * (hv - tg) * sc;
        mx = fmax(mx, hv);
    }
    part[k] = s;
    if (hmax) hmax[k] = mx;
}

template <typename T>
__global__ void k_loss_height_grad(int B, int Tw, int steps, int jp, const T *h, const T *th, T g,
                                   T *hg) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T sc = g * T(200) / T(B) / T(steps / jp);
    for (int t = jp - 1; t <= steps && t < Tw; t += jp) {
        if (t < 1) continue;
        const T tg = th[(size_t)t * B + k];
        if (tg > T(0.1) + T(1e-6)) hg[(size_t)t * B + k] += (h[(size_t)t * B + k] - tg) * sc;
    }
}

// compute_loss_pose (taichi_env.py:224-233): t in [1, steps], t % jp == 0,
// sum_i |x_i - obj_i|^2 / B / (steps / jp).  x [T][B][3][N], obj [N][3];
// block per robot.
template <typename T>
__global__ void k_loss_pose(int B, int N, int Tw, int steps, int jp, const T *x, const T *obj,
                            T *part) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *red = reinterpret_cast<T *>(smem_raw);
    const int k = blockIdx.x;
    T s[1] = {T(0)};
    for (int t = jp; t <= steps; t += jp) {
        for (int i = threadIdx.x; i < N; i += blockDim.x)
            for (int d = 0; d < 3; ++d) {
                const T xv = t < Tw ? x[((size_t)t * B + k) * 3 * N + d * N + i] : T(0);
                const T e = xv - obj[i * 3 + d];
                s[0] += e * e;
            }
    }
    block_sum<T, 1>(s, red);
    if (threadIdx.x == 0) part[k] = s[0] / T(B) / T(steps / jp);
}

template <typename T>
__global__ void k_loss_pose_grad(int B, int N, int Tw, int steps, int jp, const T *x,
                                 const T *obj, T g, T *xg) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * N) return;
    const int k = id / N, i = id - k * N;
    const T sc = g * T(2) / T(B) / T(steps / jp);
    for (int t = jp; t <= steps && t < Tw; t += jp)
        for (int d = 0; d < 3; ++d) {
            const size_t o = ((size_t)t * B + k) * 3 * N + d * N + i;
            xg[o] += (x[o] - obj[i * 3 + d]) * sc;
        }
}

// compute_loss_actuation (taichi_env.py:276-285): t in [0, steps) with
// th < 0.1 + 1e-4, max(|a| - sqrt(|tv_x| / 0.08), 0) / O / B / steps * 10 per
// output (the reference's outputs are its springs).  act [T][B][O],
// tv [T][B][3].  The adjoint of |a| is sign(a), of max(u, 0) the step u > 0.
template <typename T>
__global__ void k_loss_actuation(int B, int O, int Tw, int steps, const T *a, const T *tv,
                                 const T *th, T *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    T s = T(0);
    for (int t = 0; t < steps; ++t) {
        if (!(th[(size_t)t * B + k] < T(0.1) + T(1e-4))) continue;
        const T lim = M<T>::sqrt(fabs(tv[((size_t)t * B + k) * 3]) / T(0.08));
        for (int o = 0; o < O; ++o) {
            const T av = t < Tw ? a[((size_t)t * B + k) * O + o] : T(0);
            s += fmax(fabs(av) - lim, T(0));
        }
    }
    part[k] = s / T(O) / T(B) / T(steps) * T(10);
}

template <typename T>
__global__ void k_loss_actuation_grad(int B, int O, int Tw, int steps, const T *a, const T *tv,
                                      const T *th, T g, T *ag) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * O) return;
    const int k = id / O, o = id - k * O;
    const T sc = g / T(O) / T(B) / T(steps) * T(10);
    for (int t = 0; t < steps && t < Tw; ++t) {
        if (!(th[(size_t)t * B + k] < T(0.1) + T(1e-4))) continue;
        const T lim = M<T>::sqrt(fabs(tv[((size_t)t * B + k) * 3]) / T(0.08));
        const size_t q = ((size_t)t * B + k) * O + o;
        const T av = a[q];
        if (fabs(av) - lim > T(0)) ag[q] += (av > T(0) ? sc : (av < T(0) ? -sc : T(0)));
    }
}

// compute_loss_crawl (taichi_env.py:287-293): t in [1, steps] with
// tc > 1 - 1e-4, max(upper - 0.1, 0)^2 / B / steps * 5.
template <typename T>
__global__ void k_loss_crawl(int B, int Tw, int steps, const T *up, const T *tc, T *part) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    T s = T(0);
    for (int t = 1; t <= steps; ++t) {
        if (!(tc[(size_t)t * B + k] > T(1) - T(1e-4))) continue;
        const T u = fmax((t < Tw ? up[(size_t)t * B + k] : T(0)) - T(0.1), T(0));
        s += u * u;
    }
    part[k] = s / T(B) / T(steps) * T(5);
}

template <typename T>
__global__ void k_loss_crawl_grad(int B, int Tw, int steps, const T *up, const T *tc, T g,
                                  T *ug) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= B) return;
    const T sc = g * T(10) / T(B) / T(steps);
    for (int t = 1; t <= steps && t < Tw; ++t) {
        if (!(tc[(size_t)t * B + k] > T(1) - T(1e-4))) continue;
        const T u = up[(size_t)t * B + k] - T(0.1);
        if (u > T(0)) ug[(size_t)t * B + k] += u * sc;
    }
}

// Largest entry of hmax [B] (fixed order) into out[0] when it exceeds it.
template <typename T>
__global__ void k_max_update(int B, const T *hmax, T *out) {
    if (threadIdx.x != 0) return;
    T m = -INFINITY;
    for (int k = 0; k < B; ++k) m = fmax(m, hmax[k]);
    if (m > out[0]) out[0] = m;
}

}  // namespace dl

// ============================================================== C ABI
using namespace dl;

#define DL_DISPATCH(dtype, ...)                                   \
    do {                                                          \
        if ((dtype) == DL_F32) {                                  \
            using real = float;                                   \
            __VA_ARGS__;                                          \
        } else if ((dtype) == DL_F64) {                           \
            using real = double;                                  \
            __VA_ARGS__;                                          \
        } else {                                                  \
            dl::set_error("unknown dtype %d", (int)(dtype));      \
            return DL_E_ARG;                                      \
        }                                                         \
    } while (0)

static inline hipStream_t S_(void *s) { return reinterpret_cast<hipStream_t>(s); }
static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

static int check_topo(const dl_ms_topology *t) {
    DL_CHECK_ARG(t != nullptr, "topology is NULL");
    DL_CHECK_ARG(t->n_vert > 0 && t->n_spring >= 0 && t->n_act > 0, "bad topology sizes");
    DL_CHECK_ARG(t->spring_a && t->spring_b && t->spring_act && t->spring_L0 && t->spring_k &&
                     t->spring_coeff && t->vert_ptr && t->vert_idx && t->act_ptr && t->act_idx,
                 "topology has a NULL array");
    return 0;
}

namespace dl {

// Launch the MLP adjoint with workspace `work` (B*(2H+O) reals): stage 1
// writes every workspace entry it later reads, so the workspace needs no
// clearing between calls.
static int32_t mlp_backward_launch(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H,
                                   int32_t O, const void *in, const void *W1, const void *b1,
                                   const void *W2, const void *b2, const void *og, void *ing,
                                   void *W1g, void *b1g, void *W2g, void *b2g, void *work,
                                   void *stream) {
    DL_CHECK_ARG(B > 0 && I > 0 && H > 0 && O > 0 && in && W1 && b1 && W2 && b2 && og && W1g &&
                     b1g && W2g && b2g && work,
                 "dl_mlp_backward: bad arguments");
    DL_CHECK_ARG(act == DL_ACT_SIN || act == DL_ACT_TANH, "dl_mlp_backward: bad activation");
    if (dtype == DL_F32 && g_mlp_mfma && mlp_mfma_ok(I, H, O)) {
        if (g_mlp_mfma == 2) {
            const int IP = ((I + 3) & ~3) + 1;
            const size_t sh = sizeof(float) * ((size_t)MT * IP + 2 * (size_t)MT * (H + 1) +
                                               (size_t)MT * O);
            DL_CHECK_HIP(ensure_dynamic_lds((const void *)k_mlp_bwd1_mfma, sh));
            hipLaunchKernelGGL(k_mlp_bwd1_mfma, dim3((B + MT - 1) / MT), dim3(256), sh,
                               S_(stream), B, act, I, H, O, (const float *)in, (const float *)W1,
                               (const float *)b1, (const float *)W2, (const float *)b2,
                               (const float *)og, (float *)ing, (float *)work);
        } else {
            const size_t sh = sizeof(float) * (I + 3 * H + O);
            hipLaunchKernelGGL(k_mlp_bwd1<float>, dim3(B), dim3(256), sh, S_(stream), act, I, H,
                               O, (const float *)in, (const float *)W1, (const float *)b1,
                               (const float *)W2, (const float *)b2, (const float *)og,
                               (float *)ing, (float *)work);
        }
        const int ntI = (I + 15) / 16, tiles = (H / 16) * (ntI + 2) + 1;
        hipLaunchKernelGGL(k_mlp_bwd2_mfma, dim3(tiles), dim3(256), 0, S_(stream), B, I, H,
                           O, ntI, (const float *)in, (const float *)work, (float *)W1g,
                           (float *)b1g, (float *)W2g, (float *)b2g);
        DL_CHECK_LAUNCH();
        return 0;
    }
    DL_DISPATCH(dtype, {
        const size_t sh = sizeof(real) * (I + 3 * H + O);
        DL_CHECK_ARG(sh <= 65536, "dl_mlp_backward: input too wide for LDS (%d)", I);
        hipLaunchKernelGGL(k_mlp_bwd1<real>, dim3(B), dim3(256), sh, S_(stream), act, I, H, O,
                           (const real *)in, (const real *)W1, (const real *)b1,
                           (const real *)W2, (const real *)b2, (const real *)og, (real *)ing,
                           (real *)work);
        const long n = (long)H * I + (long)O * H + H + O;
        hipLaunchKernelGGL(k_mlp_bwd2<real>, dim3(nblk(n, 256)), dim3(256), 0, S_(stream), B, I,
                           H, O, (const real *)in, (const real *)work, (real *)W1g, (real *)b1g,
                           (real *)W2g, (real *)b2g);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

// Workspace of dl_mlp_backward_static: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// to a kernel (in host code the name is a stand-in, not a device address):
// it looks the array's address up once per device with hipGetSymbolAddress,
// with that device current, and passes that address.
static __device__ __align__(256) unsigned char scratch[DL_MLP_STATIC_WORK_BYTES];

constexpr int kMaxDevices = 64;
static void *g_scratch_addr[kMaxDevices];
static std::mutex g_scratch_mu;

static int32_t static_work_ptr(void **out) {
    int dev = -1;
    DL_CHECK_HIP(hipGetDevice(&dev));
    DL_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "static workspace: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr[dev] == nullptr) {
        void *p = nullptr;
        DL_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        DL_CHECK_ARG(p != nullptr, "static workspace: no device address");
        g_scratch_addr[dev] = p;
    }
    *out = g_scratch_addr[dev];
    return 0;
}

}  // namespace dl

extern "C" {

int32_t dl_ms_center(int32_t dtype, int32_t B, int32_t N, const void *x_t, const void *v_t,
                     void *center_t, void *vcom_t, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && x_t && center_t, "dl_ms_center: bad arguments");
    DL_CHECK_ARG(vcom_t == nullptr || v_t != nullptr, "dl_ms_center: vcom needs v");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_center<real>, dim3(B), dim3(256), 0, S_(stream), N,
                                          (const real *)x_t, (const real *)(vcom_t ? v_t : nullptr),
                                          (real *)center_t, (real *)vcom_t));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_center_grad(int32_t dtype, int32_t B, int32_t N, const void *cg, const void *vg,
                          void *xg, void *vgo, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && xg, "dl_ms_center_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_center_grad<real>, dim3(nblk((long)B * N, 256)),
                                          dim3(256), 0, S_(stream), B, N, (const real *)cg,
                                          (const real *)vg, (real *)xg, (real *)vgo));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_height(int32_t dtype, int32_t B, int32_t N, int32_t t, int32_t jump_period,
                     const void *x_t, const void *hprev, void *h, void *up, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && x_t && h && up && jump_period > 0, "dl_ms_height: bad arguments");
    const int reset = (t % jump_period == 0) || hprev == nullptr;
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_height<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, N, reset, (const real *)x_t,
                                          (const real *)hprev, (real *)h, (real *)up));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_height_grad(int32_t dtype, int32_t B, int32_t N, int32_t t, int32_t jump_period,
                          const void *x_t, const void *hprev, const void *hg, const void *ug,
                          void *xg, void *hpg, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && x_t && xg && jump_period > 0, "dl_ms_height_grad: bad arguments");
    const int reset = (t % jump_period == 0) || hprev == nullptr;
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_height_grad<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, N, reset, (const real *)x_t,
                                          (const real *)hprev, (const real *)hg, (const real *)ug,
                                          (real *)xg, (real *)hpg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_rotation(int32_t dtype, int32_t B, int32_t N, int32_t robot_id, const void *x_t,
                       const void *x_0, const void *c_0, void *rot, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && x_t && x_0 && c_0 && rot, "dl_ms_rotation: bad arguments");
    DL_CHECK_ARG(robot_id != 100 || N > 46, "dl_ms_rotation: robot 100 needs >= 47 vertices");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_rotation<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, N, robot_id, (const real *)x_t,
                                          (const real *)x_0, (const real *)c_0, (real *)rot));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_rotation_grad(int32_t dtype, int32_t B, int32_t N, int32_t robot_id,
                            const void *x_t, const void *x_0, const void *c_0, const void *rg,
                            void *xg, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && x_t && x_0 && c_0 && rg && xg, "dl_ms_rotation_grad: bad arguments");
    DL_CHECK_ARG(robot_id != 100 || N > 46, "dl_ms_rotation_grad: robot 100 needs >= 47 vertices");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_rotation_grad<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, N, robot_id, (const real *)x_t,
                                          (const real *)x_0, (const real *)c_0, (const real *)rg,
                                          (real *)xg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_nn_input(int32_t dtype, const dl_sensor_layout *lay, int32_t B, int32_t N, int32_t t,
                    int32_t offset, const void *x_t, const void *v_t, const void *center_t,
                    const void *tv_t, const void *th_t, const void *tc_t, void *in_t,
                    void *stream) {
    DL_CHECK_ARG(lay && B > 0 && N > 0 && x_t && center_t && tv_t && th_t && in_t,
                 "dl_nn_input: bad arguments");
    const Layout L = make_layout(*lay);
    DL_CHECK_ARG(L.n_tv >= 0 && L.n_tv <= 3 && L.run_period > 0 && (L.n_sin == 0 || L.n_sin > 0),
                 "dl_nn_input: bad layout");
    DL_CHECK_ARG(!L.vcom || v_t, "dl_nn_input: vcom layout needs v");
    DL_CHECK_ARG(L.dup_c == 0 || tc_t, "dl_nn_input: dup_c needs target_c");
    const int I = L.n_input(N);
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_nn_input<real>, dim3(nblk(I, 256), B), dim3(256),
                                          sizeof(real) * 16, S_(stream), L, N, I, t, offset,
                                          (const real *)x_t, (const real *)v_t,
                                          (const real *)center_t, (const real *)tv_t,
                                          (const real *)th_t, (const real *)tc_t, (real *)in_t));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_nn_input_grad(int32_t dtype, const dl_sensor_layout *lay, int32_t B, int32_t N,
                         const void *ing, void *xg, void *vg, void *cg, void *stream) {
    DL_CHECK_ARG(lay && B > 0 && N > 0 && ing, "dl_nn_input_grad: bad arguments");
    const Layout L = make_layout(*lay);
    const int I = L.n_input(N);
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_nn_input_grad<real>, dim3(B), dim3(256),
                                          sizeof(real) * 16, S_(stream), L, N, I,
                                          (const real *)ing, (real *)xg, (real *)vg, (real *)cg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_mlp_forward(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H, int32_t O,
                       const void *in, const void *W1, const void *b1, const void *W2,
                       const void *b2, void *hid, void *out, void *stream) {
    DL_CHECK_ARG(B > 0 && I > 0 && H > 0 && O > 0 && in && W1 && b1 && W2 && b2 && out,
                 "dl_mlp_forward: bad arguments");
    DL_CHECK_ARG(act == DL_ACT_SIN || act == DL_ACT_TANH, "dl_mlp_forward: bad activation");
    if (dtype == DL_F32 && g_mlp_mfma && mlp_mfma_ok(I, H, O)) {
        const int IP = ((I + 3) & ~3) + 1;
        const size_t sh = sizeof(float) * ((size_t)MT * IP + (size_t)MT * (H + 1));
        DL_CHECK_HIP(ensure_dynamic_lds((const void *)k_mlp_fwd_mfma, sh));
        hipLaunchKernelGGL(k_mlp_fwd_mfma, dim3((B + MT - 1) / MT), dim3(256), sh, S_(stream), B,
                           act, I, H, O, (const float *)in, (const float *)W1, (const float *)b1,
                           (const float *)W2, (const float *)b2, (float *)hid, (float *)out);
        DL_CHECK_LAUNCH();
        return 0;
    }
    DL_DISPATCH(dtype, {
        const size_t sh = sizeof(real) * (I + H);
        DL_CHECK_ARG(sh <= 65536, "dl_mlp_forward: input too wide for LDS (%d)", I);
        hipLaunchKernelGGL(k_mlp_fwd<real>, dim3(B), dim3(256), sh, S_(stream), act, I, H, O,
                           (const real *)in, (const real *)W1, (const real *)b1,
                           (const real *)W2, (const real *)b2, (real *)hid, (real *)out,
                           (real *)nullptr, (real *)nullptr);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_set_option(int32_t option, int32_t value) {
    DL_CHECK_ARG(option == DL_OPT_MLP_MFMA, "dl_set_option: unknown option %d", (int)option);
    DL_CHECK_ARG(value >= 0 && value <= 2, "dl_set_option: DL_OPT_MLP_MFMA takes 0, 1 or 2");
    g_mlp_mfma = value;
    return 0;
}

int32_t dl_mlp_backward(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H, int32_t O,
                           const void *in, const void *W1, const void *b1, const void *W2,
                           const void *b2, const void *og, void *ing, void *W1g, void *b1g,
                           void *W2g, void *b2g, void *work, void *stream) {
    DL_CHECK_ARG(work != nullptr, "dl_mlp_backward: bad arguments (work is NULL)");
    return dl::mlp_backward_launch(dtype, act, B, I, H, O, in, W1, b1, W2, b2, og, ing, W1g, b1g,
                                   W2g, b2g, work, stream);
}

int32_t dl_mlp_backward_static(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H,
                               int32_t O, const void *in, const void *W1, const void *b1,
                               const void *W2, const void *b2, const void *og, void *ing,
                               void *W1g, void *b1g, void *W2g, void *b2g, void *stream) {
    DL_CHECK_ARG(dtype == DL_F32 || dtype == DL_F64, "dl_mlp_backward_static: unknown dtype %d",
                 (int)dtype);
    DL_CHECK_ARG(B > 0 && H > 0 && O > 0, "dl_mlp_backward_static: bad arguments");
    const size_t need = (size_t)B * (2 * (size_t)H + (size_t)O) * (dtype == DL_F32 ? 4 : 8);
    DL_CHECK_ARG(need <= (size_t)DL_MLP_STATIC_WORK_BYTES,
                 "dl_mlp_backward_static: workspace of %zu bytes exceeds the static limit of %d "
                 "bytes (B*(2H+O) reals); use dl_mlp_backward with a caller workspace",
                 need, (int)DL_MLP_STATIC_WORK_BYTES);
    void *work = nullptr;
    if (int e = dl::static_work_ptr(&work)) return e;
    return dl::mlp_backward_launch(dtype, act, B, I, H, O, in, W1, b1, W2, b2, og, ing, W1g, b1g,
                                   W2g, b2g, work, stream);
}

int32_t dl_ms_advance(int32_t dtype, const dl_ms_topology *topo, const dl_ms_physics *phys,
                      int32_t B, const void *x_t, const void *v_t, const void *act_t, void *x_t1,
                      void *v_t1, void *stream) {
    if (int e = check_topo(topo)) return e;
    DL_CHECK_ARG(phys && B > 0 && x_t && v_t && act_t && x_t1 && v_t1, "dl_ms_advance: bad arguments");
    DL_CHECK_ARG(x_t != x_t1 && v_t != v_t1, "dl_ms_advance: in-place update not supported");
    DL_DISPATCH(dtype, {
        auto tp = make_topo<real>(*topo);
        auto P = make_phys<real>(*phys);
        hipLaunchKernelGGL(k_advance<real>, dim3(nblk((long)B * tp.N, 128)), dim3(128), 0,
                           S_(stream), tp, P, B, (const real *)x_t, (const real *)v_t,
                           (const real *)act_t, (real *)x_t1, (real *)v_t1);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_ms_advance_grad(int32_t dtype, const dl_ms_topology *topo, const dl_ms_physics *phys,
                           int32_t B, const void *x_t, const void *v_t, const void *act_t,
                           const void *xg1, const void *vg1, void *xg, void *vg, void *ag,
                           void *stream) {
    if (int e = check_topo(topo)) return e;
    DL_CHECK_ARG(phys && B > 0 && x_t && v_t && act_t && xg1 && vg1 && xg && vg,
                 "dl_ms_advance_grad: bad arguments");
    DL_DISPATCH(dtype, {
        auto tp = make_topo<real>(*topo);
        auto P = make_phys<real>(*phys);
        const size_t sh = sizeof(real) * (3 * (size_t)tp.N + 7 * (size_t)tp.S);
        DL_CHECK_ARG(sh <= 160 * 1024, "dl_ms_advance_grad: robot too large for LDS (N=%d S=%d)",
                     tp.N, tp.S);
        hipLaunchKernelGGL(k_advance_grad<real>, dim3(B), dim3(256), sh, S_(stream), tp, P,
                           (const real *)x_t, (const real *)v_t, (const real *)act_t,
                           (const real *)xg1, (const real *)vg1, (real *)xg, (real *)vg,
                           (real *)ag);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_velocity(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps,
                         int32_t rp, int32_t tp, const void *c, const void *tv, void *loss,
                         void *lb, void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && rp > 0 && tp > 0 && c && tv && loss,
                 "dl_loss_velocity: bad arguments");
    DL_CHECK_ARG(steps - rp < T, "dl_loss_velocity: targets too short");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_velocity<real>, dim3(1), dim3(256),
                                          sizeof(real) * 8, S_(stream), B, Tw, steps, rp, tp,
                                          (const real *)c, (const real *)tv, (real *)loss,
                                          (real *)lb));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_velocity_grad(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps,
                              int32_t rp, int32_t tp, const void *c, const void *tv, double g,
                              void *cg, void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && rp > 0 && tp > 0 && c && tv && cg,
                 "dl_loss_velocity_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_velocity_grad<real>, dim3(nblk(B, 64)), dim3(64),
                                          0, S_(stream), B, Tw, steps, rp, tp, (const real *)c,
                                          (const real *)tv, (real)g, (real *)cg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_rotation(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t Tw,
                            int32_t steps, int32_t rp, int32_t tp, const void *x, const void *c,
                            const void *th, void *loss, void *lb, void *work, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && T > 0 && Tw <= T && steps > 0 && rp > 0 && tp > 0 && x && c &&
                     th && loss && work,
                 "dl_loss_rotation: bad arguments");
    DL_DISPATCH(dtype, {
        hipLaunchKernelGGL(k_loss_rotation<real>, dim3(B), dim3(64), sizeof(real) * 8,
                           S_(stream), B, N, Tw, min(Tw + 1, T), steps, rp, tp, (const real *)x,
                           (const real *)c,
                           (const real *)th, (real *)work);
        hipLaunchKernelGGL(k_loss_sum<real>, dim3(1), dim3(256), sizeof(real) * 8, S_(stream), B,
                           (const real *)work, (real *)loss, (real *)lb);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_rotation_grad(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t Tw,
                              int32_t steps, int32_t rp, int32_t tp, const void *x, const void *c,
                              const void *th, double g, void *xg, void *cg, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && T > 0 && Tw <= T && steps > 0 && rp > 0 && tp > 0 && x && c &&
                     th && xg && cg,
                 "dl_loss_rotation_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_rotation_grad<real>, dim3(B), dim3(64),
                                          sizeof(real) * 24, S_(stream), B, N, Tw,
                                          min(Tw + 1, T), steps, rp, tp,
                                          (const real *)x, (const real *)c, (const real *)th,
                                          (real)g, (real *)xg, (real *)cg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_task(int32_t dtype, int32_t B, int32_t Tn, int32_t W, const void *c,
                     const void *vc, const void *tv, const void *th, void *loss, void *lb,
                     void *stream) {
    DL_CHECK_ARG(B > 0 && Tn > 0 && W > 0 && W <= Tn + 1 && c && vc && tv && th && loss,
                 "dl_loss_task: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_task<real>, dim3(1), dim3(256), sizeof(real) * 8,
                                          S_(stream), B, Tn, W, (const real *)c, (const real *)vc,
                                          (const real *)tv, (const real *)th, (real *)loss,
                                          (real *)lb));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_task_grad(int32_t dtype, int32_t B, int32_t Tn, int32_t W, const void *c,
                          const void *vc, const void *tv, const void *th, double g, void *cg,
                          void *vcg, void *stream) {
    DL_CHECK_ARG(B > 0 && Tn > 0 && W > 0 && W <= Tn + 1 && c && vc && tv && th && cg && vcg,
                 "dl_loss_task_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_task_grad<real>, dim3(nblk((long)B * W, 256)),
                                          dim3(256), 0, S_(stream), B, Tn, W, (const real *)c,
                                          (const real *)vc, (const real *)tv, (const real *)th,
                                          (real)g, (real *)cg, (real *)vcg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_height(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps, int32_t jp,
                       const void *h, const void *th, void *loss, void *lb, void *max_height,
                       void *work, void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && jp > 0 && steps / jp > 0 &&
                     steps < T && h && th && loss && work,
                 "dl_loss_height: bad arguments");
    DL_DISPATCH(dtype, {
        real *part = (real *)work, *hm = part + B;
        hipLaunchKernelGGL(k_loss_height<real>, dim3(nblk(B, 64)), dim3(64), 0, S_(stream), B, Tw,
                           steps, jp, (const real *)h, (const real *)th, part, hm);
        hipLaunchKernelGGL(k_loss_sum<real>, dim3(1), dim3(256), sizeof(real) * 8, S_(stream), B,
                           (const real *)part, (real *)loss, (real *)lb);
        if (max_height)
            hipLaunchKernelGGL(k_max_update<real>, dim3(1), dim3(64), 0, S_(stream), B,
                               (const real *)hm, (real *)max_height);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_height_grad(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps,
                            int32_t jp, const void *h, const void *th, double g, void *hg,
                            void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && jp > 0 && steps / jp > 0 &&
                     steps < T && h && th && hg,
                 "dl_loss_height_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_height_grad<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, Tw, steps, jp, (const real *)h,
                                          (const real *)th, (real)g, (real *)hg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_pose(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t Tw, int32_t steps,
                     int32_t jp, const void *x, const void *obj, void *loss, void *lb, void *work,
                     void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && T > 0 && Tw <= T && steps > 0 && jp > 0 && steps / jp > 0 &&
                     x && obj && loss && work,
                 "dl_loss_pose: bad arguments");
    DL_DISPATCH(dtype, {
        hipLaunchKernelGGL(k_loss_pose<real>, dim3(B), dim3(64), sizeof(real) * 8, S_(stream), B,
                           N, Tw, steps, jp, (const real *)x, (const real *)obj, (real *)work);
        hipLaunchKernelGGL(k_loss_sum<real>, dim3(1), dim3(256), sizeof(real) * 8, S_(stream), B,
                           (const real *)work, (real *)loss, (real *)lb);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_pose_grad(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t Tw,
                          int32_t steps, int32_t jp, const void *x, const void *obj, double g,
                          void *xg, void *stream) {
    DL_CHECK_ARG(B > 0 && N > 0 && T > 0 && Tw <= T && steps > 0 && jp > 0 && steps / jp > 0 &&
                     x && obj && xg,
                 "dl_loss_pose_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_pose_grad<real>, dim3(nblk((long)B * N, 256)),
                                          dim3(256), 0, S_(stream), B, N, Tw, steps, jp,
                                          (const real *)x, (const real *)obj, (real)g,
                                          (real *)xg));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_actuation(int32_t dtype, int32_t B, int32_t O, int32_t T, int32_t Tw,
                          int32_t steps, const void *a, const void *tv, const void *th,
                          void *loss, void *lb, void *work, void *stream) {
    DL_CHECK_ARG(B > 0 && O > 0 && T > 0 && Tw <= T && steps > 0 && steps <= T && a && tv && th &&
                     loss && work,
                 "dl_loss_actuation: bad arguments");
    DL_DISPATCH(dtype, {
        hipLaunchKernelGGL(k_loss_actuation<real>, dim3(nblk(B, 64)), dim3(64), 0, S_(stream), B,
                           O, Tw, steps, (const real *)a, (const real *)tv, (const real *)th,
                           (real *)work);
        hipLaunchKernelGGL(k_loss_sum<real>, dim3(1), dim3(256), sizeof(real) * 8, S_(stream), B,
                           (const real *)work, (real *)loss, (real *)lb);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_actuation_grad(int32_t dtype, int32_t B, int32_t O, int32_t T, int32_t Tw,
                               int32_t steps, const void *a, const void *tv, const void *th,
                               double g, void *ag, void *stream) {
    DL_CHECK_ARG(B > 0 && O > 0 && T > 0 && Tw <= T && steps > 0 && steps <= T && a && tv && th &&
                     ag,
                 "dl_loss_actuation_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_actuation_grad<real>,
                                          dim3(nblk((long)B * O, 256)), dim3(256), 0, S_(stream),
                                          B, O, Tw, steps, (const real *)a, (const real *)tv,
                                          (const real *)th, (real)g, (real *)ag));
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_crawl(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps,
                      const void *up, const void *tc, void *loss, void *lb, void *work,
                      void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && steps < T && up && tc && loss && work,
                 "dl_loss_crawl: bad arguments");
    DL_DISPATCH(dtype, {
        hipLaunchKernelGGL(k_loss_crawl<real>, dim3(nblk(B, 64)), dim3(64), 0, S_(stream), B, Tw,
                           steps, (const real *)up, (const real *)tc, (real *)work);
        hipLaunchKernelGGL(k_loss_sum<real>, dim3(1), dim3(256), sizeof(real) * 8, S_(stream), B,
                           (const real *)work, (real *)loss, (real *)lb);
    });
    DL_CHECK_LAUNCH();
    return 0;
}

int32_t dl_loss_crawl_grad(int32_t dtype, int32_t B, int32_t T, int32_t Tw, int32_t steps,
                           const void *up, const void *tc, double g, void *ug, void *stream) {
    DL_CHECK_ARG(B > 0 && T > 0 && Tw <= T && steps > 0 && steps < T && up && tc && ug,
                 "dl_loss_crawl_grad: bad arguments");
    DL_DISPATCH(dtype, hipLaunchKernelGGL(k_loss_crawl_grad<real>, dim3(nblk(B, 64)), dim3(64), 0,
                                          S_(stream), B, Tw, steps, (const real *)up,
                                          (const real *)tc, (real)g, (real *)ug));
    DL_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
