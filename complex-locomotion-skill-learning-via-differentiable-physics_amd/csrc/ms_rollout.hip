// Persistent batched mass-spring rollout: one workgroup owns one robot and
// advances it through many steps with its state resident in LDS.  Robots of
// different size, design and horizon share one launch through the offset
// table (dl_robot).  The backward kernel walks each robot's checkpoint
// segments in reverse: recompute the segment forward from its checkpoint into
// a per-robot scratch buffer, then run the hand-written adjoint back through
// it, and fold the segment's MLP weight gradients in once per segment.
//
// Per step and robot (forward):
//   COM (block reduction) -> sensors -> MLP (2 layers) -> spring impulses
//   -> per-vertex gather (CSR, fixed order) + integrate.

#include "dl_common.hpp"
#include <type_traits>
#include <utility>

namespace dl {

constexpr int HMAX = 64;

// Instruction classes of __builtin_amdgcn_sched_group_barrier (LLVM AMDGPU):
// the wave kernels pin "all LDS reads, then the arithmetic" in their LDS-bound
// inner loops, which the compiler otherwise serializes under register pressure.
constexpr int kSchedValu = 0x002, kSchedDsRead = 0x100, kSchedRead = 0x120;   // DS or VMEM reads

template <typename T>
struct RArgs {
    const dl_robot *robots;
    const int32_t *sa, *sb, *sact;
    const T *L0, *k, *coef;
    const int32_t *vptr, *vidx, *aptr, *aidx;
    const T *W;
    T *Wg;
    T *state, *adj, *ckpt, *scratch;
    const T *targets;
    T *loss;
    const T *loss_grad;
    T *traj;
    const T *traj_grad;
    int32_t *err;      // optional: bits of contract violations (DL_ROLLOUT_ERR_*)
    Phys<T> P;
    Layout L;
    int H, K, max_steps, act, loss_kind, window, store_traj;
    int max_n, max_s, max_i, max_o;
    int springs_in_lds;   // per-step spring buffers in LDS (else scratch)
    int topo_in_lds;      // design topology staged in LDS at launch start
    int csr_in_lds;       // vertex / actuator spring lists in LDS (with the topology, or
                          // alone when only they fit)
    int w_in_lds;         // W1 staged in LDS as [I][H+1] (serves W1 x and W1^T x)
};

// Per-robot view (all pointers resolved from the offset table).
template <typename T>
struct Rob {
    int N, S, O, I, Th;
    const int32_t *sa, *sb, *sact, *vptr, *vidx, *aptr, *aidx;
    const T *L0, *k, *coef;
    const T *W1, *W1T, *b1, *W2, *W2T, *b2;
    const T *WL;       // LDS copy of W1^T with row pitch H+1 (nullptr when not staged)
    T *Wg1, *bg1, *Wg2, *bg2;
    const T *tgt;
    int tgt_stride;
    T *state, *adj, *ckpt;
    T *traj;
    const T *trajg;
    // scratch: rows of a segment of K steps (buffer 0), the spring spill
    // area, and a second buffer of K rows (scr2) for the pipelined adjoint
    T *sx, *svi, *sin, *shid, *sdhid, *sout, *sdout, *scv, *sz1, *sz2, *spill, *scr2;
};

// Point the scratch rows of q at a buffer of K rows starting at s.
template <typename T>
__device__ __forceinline__ void carve_scratch(Rob<T> &q, T *s, int K, int H) {
    const int N = q.N, I = q.I, O = q.O;
    q.sx = s; s += (size_t)K * 6 * N;
    q.svi = s; s += (size_t)K * 3 * N;
    q.sin = s; s += (size_t)K * I;
    q.shid = s; s += (size_t)K * H;
    q.sdhid = s; s += (size_t)K * H;
    q.sz1 = s; s += (size_t)K * H;
    q.sout = s; s += (size_t)K * O;
    q.sdout = s; s += (size_t)K * O;
    q.sz2 = s; s += (size_t)K * O;
    q.scv = s;
}

template <typename T>
__device__ __forceinline__ Rob<T> make_rob(const RArgs<T> &A, int r) {
    const dl_robot R = A.robots[r];
    Rob<T> q;
    q.N = R.n_vert; q.S = R.n_spring; q.O = R.n_act; q.I = R.n_in; q.Th = R.horizon;
    q.sa = A.sa + R.spring0; q.sb = A.sb + R.spring0; q.sact = A.sact + R.spring0;
    q.L0 = A.L0 + R.spring0; q.k = A.k + R.spring0; q.coef = A.coef + R.spring0;
    q.vptr = A.vptr + R.vptr0; q.vidx = A.vidx;   // vptr rows hold absolute vidx positions
    q.aptr = A.aptr + R.aptr0; q.aidx = A.aidx;
    const int H = A.H, I = q.I, O = q.O;
    const T *w = A.W + R.weights;
    q.W1 = w; q.W1T = w + (size_t)H * I; q.b1 = q.W1T + (size_t)I * H;
    q.W2 = q.b1 + H; q.W2T = q.W2 + (size_t)O * H; q.b2 = q.W2T + (size_t)H * O;
    q.WL = nullptr;
    T *g = A.Wg + R.wgrad;
    q.Wg1 = g; q.bg1 = g + (size_t)H * I; q.Wg2 = q.bg1 + H; q.bg2 = q.Wg2 + (size_t)O * H;
    q.tgt = A.targets + R.targets;
    q.tgt_stride = R.target_stride;
    q.state = A.state + R.state;
    q.adj = A.adj ? A.adj + R.state : nullptr;
    q.ckpt = A.ckpt + R.ckpt;
    q.traj = (A.traj && R.traj >= 0) ? A.traj + R.traj : nullptr;
    q.trajg = (A.traj_grad && R.traj >= 0) ? A.traj_grad + R.traj : nullptr;
    const int K = A.K, N = q.N;
    T *s = A.scratch ? A.scratch + R.scratch : nullptr;
    if (s) {
        carve_scratch(q, s, K, H);
        q.spill = q.scv + (size_t)K * 6;
        q.scr2 = q.spill + (size_t)10 * q.S + 4;
        // the spilled spring rows are 16-byte rows (32 bytes in FP64): the
        // area (10 S + 4 reals) holds 8 S of them after aligning its start
        const uintptr_t al = 4 * sizeof(T);
        q.spill = reinterpret_cast<T *>((reinterpret_cast<uintptr_t>(q.spill) + al - 1) / al * al);
    } else {
        q.sx = q.svi = q.sin = q.shid = q.sdhid = q.sz1 = q.sout = q.sdout = q.sz2 = q.scv = q.spill = nullptr;
        q.scr2 = nullptr;
    }
    return q;
}

// Shared-memory carve-up (sized for the largest robot of the launch).
template <typename T>
struct Sm {
    T *xs, *vs;        // [3][N] each, stride N
    T *xb, *vb, *vib;  // adjoint state / v_inc adjoint
    T *J;              // [3][S] impulses; in the adjoint reused for [7][S] terms
    T *inb, *inbar;    // [I]
    T *hid, *dhid, *z1b;   // [HMAX]
    T *outb, *dout, *z2b;  // [O]
    T *red;            // reductions: 4*HMAX + 96
    T *cv;             // center[3], vcom[3], cbar[3], vcbar[3]
    T *topo_r;         // staged topology reals [3][S]
    T *wl;             // staged W1^T [I][H+1]
    int *topo_i;       // staged topology ints
};

template <typename T>
__device__ __forceinline__ Sm<T> carve(const RArgs<T> &A, unsigned char *raw, T *spill, bool bwd) {
    Sm<T> m;
    T *p = reinterpret_cast<T *>(raw);
    const int n3 = 3 * A.max_n;
    m.xs = p; p += n3;
    m.vs = p; p += n3;
    if (bwd) {
        m.xb = p; p += n3;
        m.vb = p; p += n3;
        m.vib = p; p += n3;
    } else {
        m.xb = m.vb = m.vib = nullptr;
    }
    if (A.springs_in_lds) {
        m.J = p;
        p += (bwd ? 7 : 3) * A.max_s;
    } else {
        m.J = spill;
    }
    m.inb = p; p += A.max_i;
    m.inbar = p; p += bwd ? A.max_i : 0;
    m.hid = p; p += HMAX;
    m.dhid = p; p += HMAX;
    m.z1b = p; p += HMAX;
    m.outb = p; p += A.max_o;
    m.dout = p; p += A.max_o;
    m.z2b = p; p += A.max_o;
    m.cv = p; p += 16;
    m.red = p; p += 16 * HMAX + 96;  // MLP partials (block <= 1024) and block sums
    m.topo_r = p; p += A.topo_in_lds ? 3 * A.max_s : 0;
    m.wl = p; p += A.w_in_lds ? (size_t)A.max_i * (HMAX + 1) : 0;
    m.topo_i = reinterpret_cast<int *>(p);
    return m;
}

template <typename T>
__host__ size_t smem_bytes(int max_n, int max_s, int max_i, int max_o, bool bwd, bool springs,
                           bool topo, bool wl, bool csr = false) {
    size_t n = 6 * (size_t)max_n + (bwd ? 9 * (size_t)max_n : 0);
    if (springs) n += (bwd ? 7 : 3) * (size_t)max_s;
    n += (bwd ? 2 : 1) * (size_t)max_i + 3 * HMAX + 3 * (size_t)max_o + 16 + 16 * HMAX + 96;
    if (topo) n += 3 * (size_t)max_s;
    if (wl) n += (size_t)max_i * (HMAX + 1);
    size_t b = n * sizeof(T);
    if (topo) b += sizeof(int) * (6 * (size_t)max_s + max_n + max_o + 2);
    else if (csr) b += sizeof(int) * (3 * (size_t)max_s + max_n + max_o + 2);
    return b;
}

// Per-spring rows of the step: impulses J (3 reals) in the forward and the
// spring adjoint terms G (dL/d dx, dL/d dv, dL/d a: 7 reals) in the adjoint.
// In LDS they are channel-major, [c][S].  Spilled to global scratch (robots
// whose springs do not fit in LDS) they are spring-major with 16-byte rows
// (J: 4 reals per spring, G: 8), so a vertex's gather costs one (J) or two
// (G) vector loads per incident spring instead of 3 or 6 scattered ones.
//
// Placement: the topology arrays (sa, sb, sact, L0, k, coef, vptr, vidx,
// aptr, aidx) are in LDS when staged and in global memory otherwise, and the
// spring rows in LDS or spilled; decided per launch.  Through a generic
// pointer every access would be a flat instruction (which waits on both the
// vector-memory and the LDS counters), so the spring passes and gathers are
// instantiated per placement -- ROWS_: rows spilled to global memory, TL:
// topology in LDS -- with address-space-qualified pointers, and the caller
// picks the instance once per pass (DL_PLACEMENT).
#define DL_AS_LDS __attribute__((address_space(3)))
#define DL_AS_GLB __attribute__((address_space(1)))
template <bool LDS_, typename T> struct AsPtr { using type = DL_AS_GLB T *; };
template <typename T> struct AsPtr<true, T> { using type = DL_AS_LDS T *; };
template <bool LDS_, typename T>
__device__ __forceinline__ typename AsPtr<LDS_, T>::type as_ptr(T *p) {
    return (typename AsPtr<LDS_, T>::type)p;
}

template <typename T> struct V4 { T a, b, c, d; };
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ V4<T> ld_v4(const T *p) {
    if constexpr (std::is_same_v<T, float>) {
        const f32x4 v = *as_ptr<false>(reinterpret_cast<const f32x4 *>(p));
        return {v.x, v.y, v.z, v.w};
    } else {
        const f64x2 u = *as_ptr<false>(reinterpret_cast<const f64x2 *>(p));
        const f64x2 v = *as_ptr<false>(reinterpret_cast<const f64x2 *>(p + 2));
        return {u.x, u.y, v.x, v.y};
    }
}
template <typename T>
__device__ __forceinline__ void st_v4(T *p, T a, T b, T c, T d) {
    if constexpr (std::is_same_v<T, float>) {
        *as_ptr<false>(reinterpret_cast<f32x4 *>(p)) = f32x4{a, b, c, d};
    } else {
        *as_ptr<false>(reinterpret_cast<f64x2 *>(p)) = f64x2{a, b};
        *as_ptr<false>(reinterpret_cast<f64x2 *>(p + 2)) = f64x2{c, d};
    }
}
template <bool ROWS_, typename T>
__device__ __forceinline__ void j_store(T *J, int S, int s, const T (&j)[3]) {
    if constexpr (ROWS_) {
        st_v4(J + 4 * (size_t)s, j[0], j[1], j[2], T(0));
    } else {
        const auto L = as_ptr<true>(J);
        L[s] = j[0]; L[S + s] = j[1]; L[2 * S + s] = j[2];
    }
}
template <bool ROWS_, typename T>
__device__ __forceinline__ void j_load(const T *J, int S, int s, T (&j)[3]) {
    if constexpr (ROWS_) {
        const V4<T> v = ld_v4(J + 4 * (size_t)s);
        j[0] = v.a; j[1] = v.b; j[2] = v.c;
    } else {
        const auto L = as_ptr<true>(J);
        j[0] = L[s]; j[1] = L[S + s]; j[2] = L[2 * S + s];
    }
}
template <bool ROWS_, typename T>
__device__ __forceinline__ void g_store(T *G, int S, int s, const T (&gd)[3], const T (&gw)[3],
                                        T ga) {
    if constexpr (ROWS_) {
        st_v4(G + 8 * (size_t)s, gd[0], gd[1], gd[2], gw[0]);
        st_v4(G + 8 * (size_t)s + 4, gw[1], gw[2], ga, T(0));
    } else {
        const auto L = as_ptr<true>(G);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            L[c * S + s] = gd[c];
            L[(3 + c) * S + s] = gw[c];
        }
        L[6 * S + s] = ga;
    }
}
template <bool ROWS_, typename T>
__device__ __forceinline__ void g_load(const T *G, int S, int s, T (&gd)[3], T (&gw)[3]) {
    if constexpr (ROWS_) {
        const V4<T> u = ld_v4(G + 8 * (size_t)s), v = ld_v4(G + 8 * (size_t)s + 4);
        gd[0] = u.a; gd[1] = u.b; gd[2] = u.c; gw[0] = u.d; gw[1] = v.a; gw[2] = v.b;
    } else {
        const auto L = as_ptr<true>(G);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gd[c] = L[c * S + s];
            gw[c] = L[(3 + c) * S + s];
        }
    }
}
template <bool ROWS_, typename T>
__device__ __forceinline__ T g_act(const T *G, int S, int s) {
    if constexpr (ROWS_) return *as_ptr<false>(G + 8 * (size_t)s + 6);
    return as_ptr<true>(G)[6 * S + s];
}

// SPR_: spring rows in LDS; TL_: the arrays the pass indexes with are in LDS
// (the spring arrays for the spring passes, the spring lists for the gathers)
#define DL_PLACEMENT(SPR_, TL_, FN_, ...)                                                \
    do {                                                                                 \
        if (SPR_) {                                                                      \
            if (TL_) FN_<false, true>(__VA_ARGS__);                                      \
            else FN_<false, false>(__VA_ARGS__);                                         \
        } else {                                                                         \
            if (TL_) FN_<true, true>(__VA_ARGS__);                                       \
            else FN_<true, false>(__VA_ARGS__);                                          \
        }                                                                                \
    } while (0)

// Spring impulses of the resident state into the step's rows (J).
template <bool ROWS_, bool TL, typename T>
__device__ void impulse_pass(const RArgs<T> &A, const Rob<T> &R, const Sm<T> &m) {
    const int N = R.N, S = R.S, tid = threadIdx.x, BT = blockDim.x;
    const auto sa = as_ptr<TL>(R.sa), sb = as_ptr<TL>(R.sb), sact = as_ptr<TL>(R.sact);
    const auto L0 = as_ptr<TL>(R.L0), kk = as_ptr<TL>(R.k), coef = as_ptr<TL>(R.coef);
    const auto xs = as_ptr<true>(m.xs), vs = as_ptr<true>(m.vs), outb = as_ptr<true>(m.outb);
    for (int s = tid; s < S; s += BT) {
        const int a = sa[s], b = sb[s];
        const T d[3] = {xs[a] - xs[b], xs[N + a] - xs[N + b], xs[2 * N + a] - xs[2 * N + b]};
        const T w[3] = {vs[a] - vs[b], vs[N + a] - vs[N + b], vs[2 * N + a] - vs[2 * N + b]};
        T J[3];
        spring_impulse<T>(d, w, L0[s], kk[s], coef[s], outb[sact[s]], A.P, J);
        j_store<ROWS_>(m.J, S, s, J);
    }
}

// Spring adjoint terms (G) from the impulse adjoints (vib) of the step.
template <bool ROWS_, bool TL, typename T>
__device__ void impulse_adj_pass(const RArgs<T> &A, const Rob<T> &R, const Sm<T> &m) {
    const int N = R.N, S = R.S, tid = threadIdx.x, BT = blockDim.x;
    const auto sa = as_ptr<TL>(R.sa), sb = as_ptr<TL>(R.sb), sact = as_ptr<TL>(R.sact);
    const auto L0 = as_ptr<TL>(R.L0), kk = as_ptr<TL>(R.k), coef = as_ptr<TL>(R.coef);
    const auto xs = as_ptr<true>(m.xs), vs = as_ptr<true>(m.vs), vib = as_ptr<true>(m.vib);
    const auto outb = as_ptr<true>(m.outb);
    for (int s = tid; s < S; s += BT) {
        const int a = sa[s], b = sb[s];
        const T d[3] = {xs[a] - xs[b], xs[N + a] - xs[N + b], xs[2 * N + a] - xs[2 * N + b]};
        const T w[3] = {vs[a] - vs[b], vs[N + a] - vs[N + b], vs[2 * N + a] - vs[2 * N + b]};
        const T Jb[3] = {vib[b] - vib[a], vib[N + b] - vib[N + a], vib[2 * N + b] - vib[2 * N + a]};
        T gd[3], gw[3], ga;
        spring_impulse_adj<T>(d, w, L0[s], kk[s], coef[s], outb[sact[s]], A.P, Jb, gd, gw, ga);
        g_store<ROWS_>(m.J, S, s, gd, gw, ga);
    }
}

// Signed sum of the impulses of vertex i's incident springs (CSR vptr/vidx:
// entry 2 s + 1 when i is the spring's first end) and of their adjoint terms.
template <bool ROWS_, bool TL, typename T>
__device__ __forceinline__ void gather_j(const Rob<T> &R, const T *J, int S, int i, T (&vi)[3]) {
    const auto vptr = as_ptr<TL>(R.vptr), vidx = as_ptr<TL>(R.vidx);
    // four entries per pinned round: the indices, then their rows (LDS or
    // global), then the sums in list order
    const int e1 = vptr[i + 1];
    int e = vptr[i];
    for (; e + 4 <= e1; e += 4) {
        int ent[4];
        T jr[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) ent[q] = vidx[e + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) j_load<ROWS_>(J, S, ent[q] >> 1, jr[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T sg = (ent[q] & 1) ? T(1) : T(-1);
            vi[0] += sg * jr[q][0]; vi[1] += sg * jr[q][1]; vi[2] += sg * jr[q][2];
        }
        __builtin_amdgcn_sched_group_barrier(kSchedRead, 4, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 12, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedRead, ROWS_ ? 4 : 12, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 24, 0);
    }
    for (; e < e1; ++e) {
        const int ent = vidx[e], s = ent >> 1;
        const T sg = (ent & 1) ? T(1) : T(-1);
        T j[3];
        j_load<ROWS_>(J, S, s, j);
        vi[0] += sg * j[0]; vi[1] += sg * j[1]; vi[2] += sg * j[2];
    }
}
template <bool ROWS_, bool TL, typename T>
__device__ __forceinline__ void gather_g(const Rob<T> &R, const T *G, int S, int i, T (&gx)[3],
                                         T (&gv)[3]) {
    const auto vptr = as_ptr<TL>(R.vptr), vidx = as_ptr<TL>(R.vidx);
    // four entries per pinned round (indices, rows, sums in list order)
    const int e1 = vptr[i + 1];
    int e = vptr[i];
    for (; e + 4 <= e1; e += 4) {
        int ent[4];
        T gd[4][3], gw[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) ent[q] = vidx[e + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) g_load<ROWS_>(G, S, ent[q] >> 1, gd[q], gw[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T sg = (ent[q] & 1) ? T(-1) : T(1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gx[c] += sg * gd[q][c];
                gv[c] += sg * gw[q][c];
            }
        }
        __builtin_amdgcn_sched_group_barrier(kSchedRead, 4, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 12, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedRead, ROWS_ ? 8 : 24, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 48, 0);
    }
    for (; e < e1; ++e) {
        const int ent = vidx[e], s = ent >> 1;
        const T sg = (ent & 1) ? T(-1) : T(1);
        T gd[3], gw[3];
        g_load<ROWS_>(G, S, s, gd, gw);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            gx[c] += sg * gd[c];
            gv[c] += sg * gw[c];
        }
    }
}
// Sum of the actuation-adjoint terms of springs e = e0, e0 + 64, ... < e1
// of an actuator's spring list (one lane's share).
template <bool ROWS_, bool TL, typename T>
__device__ __forceinline__ T act_share(const Rob<T> &R, const T *G, int S, int e0, int e1) {
    const auto aidx = as_ptr<TL>(R.aidx);
    T ab = T(0);
#pragma unroll 4
    for (int e = e0; e < e1; e += 64) ab += g_act<ROWS_>(G, S, aidx[e]);
    return ab;
}

// Copy the robot's design topology (and W1^T when requested) into LDS and
// repoint the robot view at the copies.  CSR rows are rebased to the copies.
// CSR_: the launch may stage the spring lists alone (adjoint launches only).
template <bool CSR_ = false, typename T>
__device__ void stage(const RArgs<T> &A, Rob<T> &R, Sm<T> &m) {
    const int tid = threadIdx.x, BT = blockDim.x, S = R.S, N = R.N, O = R.O;
    if (A.topo_in_lds) {
        T *tr = m.topo_r;
        int *ti = m.topo_i;
        for (int q = tid; q < S; q += BT) {
            tr[q] = R.L0[q]; tr[S + q] = R.k[q]; tr[2 * S + q] = R.coef[q];
            ti[q] = R.sa[q]; ti[S + q] = R.sb[q]; ti[2 * S + q] = R.sact[q];
        }
        const int v0 = R.vptr[0], nv = R.vptr[N] - v0;
        int *vp = ti + 3 * S, *vx = vp + N + 1;
        for (int q = tid; q <= N; q += BT) vp[q] = R.vptr[q] - v0;
        for (int q = tid; q < nv; q += BT) vx[q] = R.vidx[v0 + q];
        const int a0 = R.aptr[0], na = R.aptr[O] - a0;
        int *ap = vx + nv, *ax = ap + O + 1;
        for (int q = tid; q <= O; q += BT) ap[q] = R.aptr[q] - a0;
        for (int q = tid; q < na; q += BT) ax[q] = R.aidx[a0 + q];
        R.L0 = tr; R.k = tr + S; R.coef = tr + 2 * S;
        R.sa = ti; R.sb = ti + S; R.sact = ti + 2 * S;
        R.vptr = vp; R.vidx = vx; R.aptr = ap; R.aidx = ax;
    } else if (CSR_ && A.csr_in_lds) {
        // only the vertex and actuator spring lists (the gathers' indices)
        const int v0 = R.vptr[0], nv = R.vptr[N] - v0;
        int *vp = m.topo_i, *vx = vp + N + 1;
        for (int q = tid; q <= N; q += BT) vp[q] = R.vptr[q] - v0;
        for (int q = tid; q < nv; q += BT) vx[q] = R.vidx[v0 + q];
        const int a0 = R.aptr[0], na = R.aptr[O] - a0;
        int *ap = vx + nv, *ax = ap + O + 1;
        for (int q = tid; q <= O; q += BT) ap[q] = R.aptr[q] - a0;
        for (int q = tid; q < na; q += BT) ax[q] = R.aidx[a0 + q];
        R.vptr = vp; R.vidx = vx; R.aptr = ap; R.aidx = ax;
    }
    if (A.w_in_lds) {
        const int H = A.H, I = R.I;
        for (int q = tid; q < I * H; q += BT) {
            const int i = q / H, j = q - i * H;
            m.wl[i * (HMAX + 1) + j] = R.W1T[q];
        }
        R.WL = m.wl;
    }
    __syncthreads();
}

// COM position and velocity of the resident state -> m.cv[0..5]
template <typename T>
__device__ __forceinline__ void com(const Rob<T> &R, Sm<T> &m) {
    T s[6] = {0, 0, 0, 0, 0, 0};
    const T invN = T(1) / T(R.N);
    for (int i = threadIdx.x; i < R.N; i += blockDim.x) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s[d] += m.xs[d * R.N + i] * invN;
            s[3 + d] += m.vs[d * R.N + i] * invN;
        }
    }
    block_sum<T, 6>(s, m.red);
    if (threadIdx.x < 6) m.cv[threadIdx.x] = s[threadIdx.x];
    __syncthreads();
}

template <typename T>
__device__ __forceinline__ const T *tgt_at(const Rob<T> &R, int t) {
    if (R.tgt_stride == 0) return R.tgt;
    const int tt = t < R.Th ? t : R.Th - 1;
    return R.tgt + (size_t)tt * R.tgt_stride;
}

// Windowed COM loss term |vc_xz - tv_xz|^2 + (c_y - th)^2
template <typename T>
__device__ __forceinline__ T task_term(const T *cv, const T *tg) {
    const T a = cv[3] - tg[0], b = cv[5] - tg[2], c = cv[1] - tg[3];
    return a * a + b * b + c * c;
}

enum { MODE_FWD = 0, MODE_REC = 1 };

// W1^T [I][64] in global memory read as 16-byte rows: FP32, hidden width 64,
// base 16-byte aligned (dl_weights_size keeps every packed block aligned).
template <typename T>
__device__ __forceinline__ bool w1_rows16(const Rob<T> &R, int H) {
    if constexpr (std::is_same_v<T, float>)
        return !R.WL && H == HMAX && (reinterpret_cast<uintptr_t>(R.W1T) & 15) == 0;
    return false;
}

// Layer-1 pre-activations z = W1 x from global W1^T (the robots whose W1 does
// not fit in LDS): thread (q = tid % 16, c = tid / 16) accumulates hidden
// units 4q..4q+3 over rows i = c, c + C, ... (C = BT / 16), one 16-byte load
// per row, so a wavefront reads 4 consecutive rows (1 KiB) per step and keeps
// eight steps in flight.  The four row groups of a wavefront are summed by
// lane shuffles, the wavefronts' partials in LDS (red: BT / 64 rows of 64);
// threads tid < 64 return z for hidden unit tid (bias included).
__device__ __forceinline__ float layer1_rows16(const float *W1T, const float *b1, const float *x,
                                               int I, float *red) {
    constexpr int U = 8;
    const int tid = threadIdx.x, q = tid & 15, g = tid >> 4, C = blockDim.x >> 4;
    const float4 *w4 = reinterpret_cast<const float4 *>(W1T) + q;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    // U rows' loads issued before their products (held in distinct
    // registers, so the loads overlap even where the kernel's register
    // budget is tight); rows past I contribute nothing
    for (int i0 = g; i0 < I; i0 += U * C) {
        float4 w[U];
        float xi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * C;
            w[u] = i < I ? w4[(size_t)i * (HMAX / 4)] : make_float4(0.f, 0.f, 0.f, 0.f);
            xi[u] = i < I ? x[i] : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0 += w[u].x * xi[u]; a1 += w[u].y * xi[u]; a2 += w[u].z * xi[u]; a3 += w[u].w * xi[u];
        }
    }
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
        a0 += __shfl_xor(a0, off, 64); a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64); a3 += __shfl_xor(a3, off, 64);
    }
    if ((tid & 63) < 16) {
        float *r = red + (tid >> 6) * HMAX + 4 * q;
        r[0] = a0; r[1] = a1; r[2] = a2; r[3] = a3;
    }
    __syncthreads();
    float z = 0.f;
    if (tid < HMAX) {
        z = b1[tid];
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) z += red[w * HMAX + tid];
    }
    return z;
}

// Input adjoint xb = W1^T z1b from global W1^T rows (as layer1_rows16): the
// 16 lanes of a DPP row take one input row (4 hidden units each), their
// products are summed by an inclusive row scan (lane 15 holds the total), and
// the workgroup's 16-lane groups stride over the rows, eight rows' loads in
// flight per lane.  The trip count is the same for every lane (rows past I
// load nothing and store nothing), so the lane shifts run with all lanes
// active.
__device__ __forceinline__ void input_adj_rows16(const float *W1T, const float *z1b, int I,
                                                 float *xb) {
    constexpr int U = 8;
    const int tid = threadIdx.x, q = tid & 15, g = tid >> 4, C = blockDim.x >> 4;
    const float4 *w4 = reinterpret_cast<const float4 *>(W1T) + q;
    const float z0 = z1b[4 * q], z1 = z1b[4 * q + 1], z2 = z1b[4 * q + 2], z3 = z1b[4 * q + 3];
    const int nit = (I + C - 1) / C;
    for (int k0 = 0; k0 < nit; k0 += U) {
        float p[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = g + (k0 + u) * C;
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i < I) w = w4[(size_t)i * (HMAX / 4)];
            p[u] = w.x * z0 + w.y * z1 + w.z * z2 + w.w * z3;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            float v = p[u];
            v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, true));
            v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xF, 0xF, true));
            v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, true));
            v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, true));
            p[u] = v;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = g + (k0 + u) * C;
            if (q == 15 && i < I) xb[i] = p[u];
        }
    }
}

// One forward step from the resident state at step t.  In MODE_REC (and in
// every mode with ROWS) the step inputs and intermediates are written to
// scratch slot `slot`.
template <typename T, bool ROWS = false>
__device__ void fwd_step(const RArgs<T> &A, const Rob<T> &R, Sm<T> &m, int t, int mode, int slot,
                         T &lsum) {
    const int N = R.N, S = R.S, I = R.I, O = R.O, H = A.H;
    const int tid = threadIdx.x, BT = blockDim.x;
    com(R, m);
    const T *tg = tgt_at(R, t);
    if (mode == MODE_FWD) {
        if (t % A.K == 0) {
            T *ck = R.ckpt + (size_t)(t / A.K) * 6 * N;
            for (int i = tid; i < 3 * N; i += BT) {
                ck[i] = m.xs[i];
                ck[3 * N + i] = m.vs[i];
            }
        }
        if (A.store_traj && R.traj) {
            T *row = R.traj + (size_t)t * (3 * N + 3);
            for (int i = tid; i < 3 * N; i += BT) row[i] = m.xs[i];
            if (tid < 3) row[3 * N + tid] = m.cv[tid];
        }
        if (A.loss_kind == 1 && t > R.Th - A.window && tid == 0) lsum += task_term(m.cv, tg);
    }
    const bool rows = ROWS || mode == MODE_REC;
    if (rows) {
        T *sx = R.sx + (size_t)slot * 6 * N;
        for (int i = tid; i < 3 * N; i += BT) {
            sx[i] = m.xs[i];
            sx[3 * N + i] = m.vs[i];
        }
        if (tid < 6) R.scv[slot * 6 + tid] = m.cv[tid];
    }
    // sensors
    for (int q = tid; q < I; q += BT) {
        const T val = sensor_value<T>(A.L, q, t, 0, N, m.xs, m.cv, m.cv + 3, tg);
        m.inb[q] = val;
        if (rows) R.sin[(size_t)slot * I + q] = val;
    }
    __syncthreads();
    // MLP layer 1: thread (j, chunk c); W1T rows are coalesced across j
    {
        T z = T(0);
        if (w1_rows16(R, H)) {
            if constexpr (std::is_same_v<T, float>) z = layer1_rows16(R.W1T, R.b1, m.inb, I, m.red);
        } else {
            const int j = tid & (HMAX - 1), c = tid / HMAX, C = BT / HMAX;
            const int Ic = (I + C - 1) / C;
            const int i0 = c * Ic, i1 = min(I, i0 + Ic);
            T acc = T(0);
            if (j < H) {
                if (R.WL) {
#pragma unroll 8
                    for (int i = i0; i < i1; ++i) acc += R.WL[i * (HMAX + 1) + j] * m.inb[i];
                } else {
#pragma unroll 8
                    for (int i = i0; i < i1; ++i) acc += R.W1T[(size_t)i * H + j] * m.inb[i];
                }
            }
            m.red[c * HMAX + j] = acc;
            __syncthreads();
            if (tid < H) {
                z = R.b1[tid];
                for (int cc = 0; cc < C; ++cc) z += m.red[cc * HMAX + tid];
            }
        }
        if (tid < H) {
            T y, dy;
            activate<T>(A.act, z, y, dy);
            m.hid[tid] = y;
            m.dhid[tid] = dy;
            if (rows) {
                R.shid[(size_t)slot * H + tid] = y;
                R.sdhid[(size_t)slot * H + tid] = dy;
            }
        }
        __syncthreads();
    }
    // MLP layer 2: P lanes per output
    {
        int P = 64;
        while (P > 1 && P * O > BT) P >>= 1;
        const int p = tid % P;
        for (int o0 = tid / P; o0 < ((O + BT / P - 1) / (BT / P)) * (BT / P); o0 += BT / P) {
            T acc = T(0);
            if (o0 < O)
                for (int jj = p; jj < H; jj += P) acc += R.W2[(size_t)o0 * H + jj] * m.hid[jj];
            for (int off = P >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if (p == 0 && o0 < O) {
                T y, dy;
                activate<T>(A.act, acc + R.b2[o0], y, dy);
                m.outb[o0] = y;
                m.dout[o0] = dy;
                if (rows) {
                    R.sout[(size_t)slot * O + o0] = y;
                    R.sdout[(size_t)slot * O + o0] = dy;
                }
            }
        }
        __syncthreads();
    }
    // spring impulses
    DL_PLACEMENT(A.springs_in_lds, A.topo_in_lds, impulse_pass, A, R, m);
    __syncthreads();
    // gather + integrate (vertex owned by one thread, in place)
    for (int i = tid; i < N; i += BT) {
        T vi[3] = {0, 0, 0};
        DL_PLACEMENT(A.springs_in_lds, A.csr_in_lds, gather_j, R, m.J, S, i, vi);
        if constexpr (ROWS) {
            // kept for the adjoint, which then skips its spring pass
            T *sv = R.svi + (size_t)slot * 3 * N;
            sv[i] = vi[0]; sv[N + i] = vi[1]; sv[2 * N + i] = vi[2];
        }
        const T x[3] = {m.xs[i], m.xs[N + i], m.xs[2 * N + i]};
        const T v[3] = {m.vs[i], m.vs[N + i], m.vs[2 * N + i]};
        T xn[3], vn[3];
        integrate<T>(x, v, vi, A.P, xn, vn);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xs[d * N + i] = xn[d];
            m.vs[d * N + i] = vn[d];
        }
    }
    __syncthreads();
}

// LB: the largest workgroup the instantiation is launched with (256, 512, or 1024 in FP32).
// ROWS: also write every step's adjoint rows (slot = step; K >= horizon).
template <typename T, int LB, bool ROWS>
__global__ void __launch_bounds__(LB) k_rollout_fwd(RArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int r = blockIdx.x;
    Rob<T> R = make_rob(A, r);
    Sm<T> m = carve(A, smem_raw, R.spill, false);
    stage(A, R, m);
    const int N = R.N, tid = threadIdx.x, BT = blockDim.x;
    const int t0 = A.robots[r].t_fwd;
    int t1 = R.Th;
    if (A.max_steps > 0 && t0 + A.max_steps < t1) t1 = t0 + A.max_steps;
    if (t0 >= t1) return;
    if (ROWS && (R.Th > A.K || R.sx == nullptr)) {
        // the stored rows need K >= horizon rows of scratch
        if (tid == 0 && A.loss) A.loss[r] = T(NAN);
        if (tid == 0 && A.err) atomicOr(A.err, DL_ROLLOUT_ERR_HORIZON);
        return;
    }
    for (int i = tid; i < 3 * N; i += BT) {
        m.xs[i] = R.state[i];
        m.vs[i] = R.state[3 * N + i];
    }
    __syncthreads();
    T lsum = T(0);
    for (int t = t0; t < t1; ++t) fwd_step<T, ROWS>(A, R, m, t, MODE_FWD, t, lsum);
    if (t1 == R.Th) {
        // final state: loss term and trajectory row (the center of the final
        // state is not part of the trajectory, matching the reference field)
        com(R, m);
        if (A.loss_kind == 1 && tid == 0) lsum += task_term(m.cv, tgt_at(R, R.Th));
        if (A.store_traj && R.traj) {
            T *row = R.traj + (size_t)R.Th * (3 * N + 3);
            for (int i = tid; i < 3 * N; i += BT) row[i] = m.xs[i];
        }
    }
    for (int i = tid; i < 3 * N; i += BT) {
        R.state[i] = m.xs[i];
        R.state[3 * N + i] = m.vs[i];
    }
    if (A.loss_kind == 1 && tid == 0) A.loss[r] += lsum / T(A.window);
}

// Adjoint of one step t.  On entry m.xb/m.vb hold dL/d(state t+1); on exit
// dL/d(state t).  Intermediates come from scratch slot `slot`; with STORED
// the step's impulse sums too (kept by the forward), so no spring pass.
template <typename T, bool STORED = false>
__device__ void bwd_step(const RArgs<T> &A, const Rob<T> &R, Sm<T> &m, int r, int t, int slot) {
    const int N = R.N, S = R.S, I = R.I, O = R.O, H = A.H;
    const int tid = threadIdx.x, BT = blockDim.x;
    const T *sx = R.sx + (size_t)slot * 6 * N;
    for (int i = tid; i < 3 * N; i += BT) {
        m.xs[i] = sx[i];
        m.vs[i] = sx[3 * N + i];
    }
    for (int q = tid; q < I; q += BT) m.inb[q] = R.sin[(size_t)slot * I + q];
    for (int j = tid; j < H; j += BT) m.dhid[j] = R.sdhid[(size_t)slot * H + j];
    for (int o = tid; o < O; o += BT) {
        m.outb[o] = R.sout[(size_t)slot * O + o];
        m.dout[o] = R.sdout[(size_t)slot * O + o];
    }
    if (tid < 6) m.cv[tid] = R.scv[slot * 6 + tid];
    __syncthreads();
    T *G = m.J;  // overwrites the impulses once the integrator adjoint consumed them
    if constexpr (!STORED) {
        // recompute impulses
        DL_PLACEMENT(A.springs_in_lds, A.topo_in_lds, impulse_pass, A, R, m);
        __syncthreads();
    }
    // integrate adjoint
    for (int i = tid; i < N; i += BT) {
        T vi[3] = {0, 0, 0};
        if constexpr (STORED) {
            const T *sv = R.svi + (size_t)slot * 3 * N;
            vi[0] = sv[i]; vi[1] = sv[N + i]; vi[2] = sv[2 * N + i];
        } else {
            DL_PLACEMENT(A.springs_in_lds, A.csr_in_lds, gather_j, R, m.J, S, i, vi);
        }
        const T x[3] = {m.xs[i], m.xs[N + i], m.xs[2 * N + i]};
        const T v[3] = {m.vs[i], m.vs[N + i], m.vs[2 * N + i]};
        const T xnb[3] = {m.xb[i], m.xb[N + i], m.xb[2 * N + i]};
        const T vnb[3] = {m.vb[i], m.vb[N + i], m.vb[2 * N + i]};
        T xb[3], vb[3], vib[3];
        integrate_adj<T>(x, v, vi, A.P, xnb, vnb, xb, vb, vib);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xb[d * N + i] = xb[d];
            m.vb[d * N + i] = vb[d];
            m.vib[d * N + i] = vib[d];
        }
    }
    __syncthreads();
    // spring adjoint
    DL_PLACEMENT(A.springs_in_lds, A.topo_in_lds, impulse_adj_pass, A, R, m);
    __syncthreads();
    for (int i = tid; i < N; i += BT) {
        T gx[3] = {0, 0, 0}, gv[3] = {0, 0, 0};
        DL_PLACEMENT(A.springs_in_lds, A.csr_in_lds, gather_g, R, G, S, i, gx, gv);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xb[d * N + i] += gx[d];
            m.vb[d * N + i] += gv[d];
        }
    }
    // actuator adjoints: one wavefront per actuator, its springs strided over
    // the lanes and summed in a fixed order (a large robot's actuator owns
    // hundreds of springs)
    {
        const int lane = tid & 63, NW = BT >> 6;
        for (int o = tid >> 6; o < O; o += NW) {
            const int e0 = R.aptr[o] + lane, e1 = R.aptr[o + 1];
            T ab = T(0);
            DL_PLACEMENT(A.springs_in_lds, A.csr_in_lds, ab = act_share, R, G, S, e0, e1);
            ab = wave_sum(ab);
            if (lane == 0) {
                const T z = ab * m.dout[o];
                m.z2b[o] = z;
                R.sz2[(size_t)slot * O + o] = z;
            }
        }
    }
    __syncthreads();
    // MLP adjoint: hidden; wavefront g sums the outputs o = g, g + NW, ...
    // (W2 rows coalesced across the hidden units), the partials in LDS
    {
        const int j = tid & (HMAX - 1), g = tid / HMAX, NW = BT / HMAX;
        T acc = T(0);
        if (j < H)
            for (int o = g; o < O; o += NW) acc += R.W2[(size_t)o * H + j] * m.z2b[o];
        m.red[g * HMAX + j] = acc;
        __syncthreads();
        if (tid < H) {
            T zs = T(0);
            for (int w = 0; w < NW; ++w) zs += m.red[w * HMAX + tid];
            const T z = zs * m.dhid[tid];
            m.z1b[tid] = z;
            R.sz1[(size_t)slot * H + tid] = z;
        }
        __syncthreads();
    }
    // input adjoint: in_b = W1^T z1b (W1 [H][I] rows coalesced across i)
    if (w1_rows16(R, H)) {
        if constexpr (std::is_same_v<T, float>) input_adj_rows16(R.W1T, m.z1b, I, m.inbar);
    } else for (int i = tid; i < I; i += BT) {
        T acc = T(0);
        if (R.WL) {
            const T *row = R.WL + i * (HMAX + 1);
#pragma unroll 16
            for (int j = 0; j < H; ++j) acc += row[j] * m.z1b[j];
        } else {
#pragma unroll 16
            for (int j = 0; j < H; ++j) acc += R.W1[(size_t)j * I + i] * m.z1b[j];
        }
        m.inbar[i] = acc;
    }
    __syncthreads();
    // sensor adjoint, loss, external trajectory adjoint
    const Layout &L = A.L;
    const int base = L.vert_base(), vs = L.vstride();
    const T inv = T(1) / T(L.offset_scale);
    T cb[3] = {0, 0, 0};
    const T *eg = R.trajg ? R.trajg + (size_t)t * (3 * N + 3) : nullptr;
    for (int i = tid; i < N; i += BT) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T g = m.inbar[base + i * vs + d] * inv;
            T add = g;
            if (eg) add += eg[d * N + i];
            m.xb[d * N + i] += add;
            cb[d] -= g;
        }
    }
    block_sum<T, 3>(cb, m.red);
    T vcb[3] = {0, 0, 0};
    if (L.vcom)
        for (int d = 0; d < 3; ++d) vcb[d] = m.inbar[L.n_sin + d];
    if (eg)
        for (int d = 0; d < 3; ++d) cb[d] += eg[3 * N + d];
    if (A.loss_kind == 1 && t > R.Th - A.window) {
        const T lg = A.loss_grad[r] / T(A.window);
        const T *tg = tgt_at(R, t);
        vcb[0] += T(2) * lg * (m.cv[3] - tg[0]);
        vcb[2] += T(2) * lg * (m.cv[5] - tg[2]);
        cb[1] += T(2) * lg * (m.cv[1] - tg[3]);
    }
    const T invN = T(1) / T(N);
    for (int i = tid; i < N; i += BT) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xb[d * N + i] += cb[d] * invN;
            m.vb[d * N + i] += vcb[d] * invN;
        }
    }
    __syncthreads();
}

// Fold the segment's weight gradients into the robot's partial:
// W1g += sum_t z1b_t in_t^T, b1g += sum_t z1b_t, W2g += sum_t z2b_t h_t^T, ...
// The step loops are unrolled so that several steps' row loads are in flight
// (each iteration's loads are independent; the sums keep the step order).
template <typename T>
__device__ void seg_wgrad(const RArgs<T> &A, const Rob<T> &R, int n, T *zs, int zcap) {
    const int I = R.I, O = R.O, H = A.H, tid = threadIdx.x, BT = blockDim.x;
    for (int i0 = 0; i0 < I; i0 += BT) {
        const int i = i0 + tid;
        const bool ok = i < I;
        T acc[HMAX];
#pragma unroll
        for (int j = 0; j < HMAX; ++j) acc[j] = T(0);
        if (H == HMAX && zcap > 0) {
            // the z rows staged in LDS, zcap steps at a time (one cooperative
            // load, then broadcast reads): read straight from global memory,
            // each guarded entry would wait for its own load
            for (int s0 = 0; s0 < n; s0 += zcap) {
                const int c = min(zcap, n - s0);
                __syncthreads();
                for (int q = tid; q < c * HMAX; q += BT) zs[q] = R.sz1[(size_t)s0 * HMAX + q];
                __syncthreads();
                for (int s = 0; s < c; ++s) {
                    const T xin = ok ? R.sin[(size_t)(s0 + s) * I + i] : T(0);
                    // 16 entries per group, the groups not interleaved by the
                    // scheduler (acc already holds 64 registers)
#pragma unroll
                    for (int j0 = 0; j0 < HMAX; j0 += 16) {
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int j = j0; j < j0 + 16; ++j) acc[j] += zs[s * HMAX + j] * xin;
                    }
                }
            }
        } else {
            for (int s = 0; s < n; ++s) {
                const T xin = ok ? R.sin[(size_t)s * I + i] : T(0);
                const T *z = R.sz1 + (size_t)s * H;
#pragma unroll
                for (int j = 0; j < HMAX; ++j)
                    if (j < H) acc[j] += z[j] * xin;
            }
        }
        if (ok && H == HMAX) {
            // read-modify-write in groups of 16 independent loads
#pragma unroll
            for (int j0 = 0; j0 < HMAX; j0 += 16) {
                T g[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) g[j] = R.Wg1[(size_t)(j0 + j) * I + i];
#pragma unroll
                for (int j = 0; j < 16; ++j) R.Wg1[(size_t)(j0 + j) * I + i] = g[j] + acc[j0 + j];
            }
        } else if (ok) {
#pragma unroll
            for (int j = 0; j < HMAX; ++j)
                if (j < H) R.Wg1[(size_t)j * I + i] += acc[j];
        }
    }
    for (int j = tid; j < H; j += BT) {
        T acc = T(0);
#pragma unroll 8
        for (int s = 0; s < n; ++s) acc += R.sz1[(size_t)s * H + j];
        R.bg1[j] += acc;
    }
    for (int q = tid; q < O * H; q += BT) {
        const int o = q / H, j = q - o * H;
        T acc = T(0);
#pragma unroll 8
        for (int s = 0; s < n; ++s) acc += R.sz2[(size_t)s * O + o] * R.shid[(size_t)s * H + j];
        R.Wg2[q] += acc;
    }
    for (int o = tid; o < O; o += BT) {
        T acc = T(0);
#pragma unroll 8
        for (int s = 0; s < n; ++s) acc += R.sz2[(size_t)s * O + o];
        R.bg2[o] += acc;
    }
}

// STORED: the forward kept every step's rows (K >= horizon): no segment
// recompute.  Otherwise each segment is recomputed from its checkpoint with
// the same rows, per-vertex impulse sums included.
template <typename T, int LB, bool STORED>
__global__ void __launch_bounds__(LB) k_rollout_bwd(RArgs<T> A) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int r = blockIdx.x;
    Rob<T> R = make_rob(A, r);
    Sm<T> m = carve(A, smem_raw, R.spill, true);
    stage<true>(A, R, m);
    const int N = R.N, tid = threadIdx.x, BT = blockDim.x, K = A.K;
    int tb = A.robots[r].t_bwd;
    if (tb <= 0) return;
    if (STORED && R.Th > K) {
        if (tid == 0) R.Wg1[0] = T(NAN);
        return;
    }
    for (int i = tid; i < 3 * N; i += BT) {
        m.xb[i] = R.adj[i];
        m.vb[i] = R.adj[3 * N + i];
    }
    __syncthreads();
    if (tb == R.Th) {
        // seed from the final state's loss term and external trajectory row
        for (int i = tid; i < 3 * N; i += BT) {
            m.xs[i] = R.state[i];
            m.vs[i] = R.state[3 * N + i];
        }
        __syncthreads();
        com(R, m);
        T cb[3] = {0, 0, 0}, vcb[3] = {0, 0, 0};
        if (A.loss_kind == 1) {
            const T lg = A.loss_grad[r] / T(A.window);
            const T *tg = tgt_at(R, R.Th);
            vcb[0] = T(2) * lg * (m.cv[3] - tg[0]);
            vcb[2] = T(2) * lg * (m.cv[5] - tg[2]);
            cb[1] = T(2) * lg * (m.cv[1] - tg[3]);
        }
        const T *eg = R.trajg ? R.trajg + (size_t)R.Th * (3 * N + 3) : nullptr;
        const T invN = T(1) / T(N);
        for (int i = tid; i < N; i += BT) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                m.xb[d * N + i] += cb[d] * invN + (eg ? eg[d * N + i] : T(0));
                m.vb[d * N + i] += vcb[d] * invN;
            }
        }
        __syncthreads();
    }
    int budget = A.max_steps > 0 ? A.max_steps : 0x7fffffff;
    T dummy = T(0);
    while (tb > 0 && budget > 0) {
        const int seg = (tb - 1) / K, s0 = seg * K;
        const int lo = max(s0, tb - budget);
        if (!STORED) {
            const T *ck = R.ckpt + (size_t)seg * 6 * N;
            for (int i = tid; i < 3 * N; i += BT) {
                m.xs[i] = ck[i];
                m.vs[i] = ck[3 * N + i];
            }
            __syncthreads();
            // the recompute keeps the impulse sums too: no spring pass below
            for (int t = s0; t < tb; ++t) fwd_step<T, true>(A, R, m, t, MODE_REC, t - s0, dummy);
        }
        for (int t = tb - 1; t >= lo; --t) bwd_step<T, true>(A, R, m, r, t, t - s0);
        __syncthreads();
        // the weight-gradient fold reads slots lo - s0 .. tb - s0 - 1
        {
            Rob<T> Rs = R;
            const int off = lo - s0;
            Rs.sin = R.sin + (size_t)off * R.I;
            Rs.sz1 = R.sz1 + (size_t)off * A.H;
            Rs.sz2 = R.sz2 + (size_t)off * R.O;
            Rs.shid = R.shid + (size_t)off * A.H;
            // the resident positions / velocities (6 max_n reals) are dead
            // until the next segment's recompute: they hold the z rows
            seg_wgrad(A, Rs, tb - lo, m.xs, 6 * A.max_n / HMAX);
        }
        __syncthreads();
        budget -= tb - lo;
        tb = lo;
    }
    for (int i = tid; i < 3 * N; i += BT) {
        R.adj[i] = m.xb[i];
        R.adj[3 * N + i] = m.vb[i];
    }
}

// ===================================================================
// Wave-per-robot rollout (FP32, MLP input width <= IMAX).
//
// A workgroup holds WAVES robots, one wavefront each, so every SIMD of a CU
// advances its own robot with no workgroup barrier inside the step loop.
// Both GEMVs read one LDS copy of W1 shared by the workgroup's robots
// (robots of one design are packed together; groups of four inputs, see
// wave_wl4_reals): the forward's lane j reads hidden unit j's weights 16
// bytes at a time, the adjoint's W1^T z has lane i read input i's weights.
// The design topology is staged into LDS once per workgroup.
// ===================================================================


__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
struct WSm {
    T *xs, *vs, *xb, *vb, *vib, *J, *inb, *inbar, *z1b, *outb, *dout, *z2b;
};

__host__ __device__ inline size_t wave_region(int max_n, int max_s, int max_i, int max_o, bool bwd) {
    const size_t ni = (size_t)(max_i > 192 ? max_i : 192) + 4;   // GEMV reads up to 192 inputs
    size_t n = (bwd ? 20 : 8) * (size_t)max_n + (bwd ? 8 : 4) * (size_t)max_s +
               (bwd ? 2 : 1) * ni + HMAX + 3 * (size_t)max_o;
    return (n + 3) & ~(size_t)3;
}

// The wave kernels' LDS copy of W1: inputs in groups of four, entry (g, j)
// = W1[j][4g .. 4g+3] at 4 (g (HMAX + 1) + j).  Hidden unit j reads one
// 16-byte entry per group (layer 1); input i reads W1[j][i] for j = 0, 1, ..
// at a stride of 4 reals (input adjoint), and the group pitch of HMAX + 1
// entries puts the 64 lanes' reads in distinct banks.  The last group padded.
__host__ __device__ inline size_t wave_wl4_reals(int max_i) {
    return (size_t)((max_i + 3) / 4) * 4 * (HMAX + 1);
}
__device__ __forceinline__ int wl4_at(int i, int j) { return 4 * ((i >> 2) * (HMAX + 1) + j) + (i & 3); }

__host__ inline size_t wave_smem_bytes(int max_n, int max_s, int max_i, int max_o, bool bwd,
                                       int n_tab, int waves) {
    size_t n = waves * wave_region(max_n, max_s, max_i, max_o, bwd) + 3 * (size_t)max_s + n_tab +
               wave_wl4_reals(max_i);
    return n * sizeof(float) + sizeof(int) * (6 * (size_t)max_s + max_n + max_o + 2);
}

// Pipelined adjoint: per robot an adjoint region and a recompute region.
__host__ inline size_t wave_pipe_smem_bytes(int max_n, int max_s, int max_i, int max_o, int n_tab,
                                            int robots) {
    size_t n = robots * (wave_region(max_n, max_s, max_i, max_o, true) +
                         wave_region(max_n, max_s, max_i, max_o, false)) +
               3 * (size_t)max_s + n_tab + wave_wl4_reals(max_i);
    return n * sizeof(float) + sizeof(int) * (6 * (size_t)max_s + max_n + max_o + 2);
}

template <typename T>
__device__ __forceinline__ WSm<T> wave_carve_at(const RArgs<T> &A, T *p, bool bwd) {
    WSm<T> m;
    const int n3 = 4 * A.max_n;   // vertex rows of 4 (x, y, z, 0)
    // spring rows first: the region starts 16-byte aligned, and the rows
    // are read and written 16 bytes at a time
    m.J = p; p += (bwd ? 8 : 4) * A.max_s;
    m.xs = p; p += n3;
    m.vs = p; p += n3;
    if (bwd) {
        m.xb = p; p += n3;
        m.vb = p; p += n3;
        m.vib = p; p += n3;
    } else {
        m.xb = m.vb = m.vib = nullptr;
    }
    const int ni = (A.max_i > 192 ? A.max_i : 192) + 4;
    m.inb = p; p += ni;
    if (bwd) { m.inbar = p; p += ni; } else { m.inbar = nullptr; }
    m.z1b = p; p += HMAX;
    m.outb = p; p += A.max_o;
    m.dout = p; p += A.max_o;
    m.z2b = p;
    return m;
}

template <typename T>
__device__ __forceinline__ WSm<T> wave_carve(const RArgs<T> &A, T *base, int wid, bool bwd) {
    return wave_carve_at(A, base + (size_t)wid * wave_region(A.max_n, A.max_s, A.max_i, A.max_o, bwd),
                         bwd);
}

// Block-wide: stage the topology of robot r0's design and its W1 (groups of
// four inputs, wave_wl4_reals; first in the shared region, so 16-byte
// aligned) into the shared region; returns pointers through `sh`.
template <typename T>
struct WShared {
    T *topo_r, *wl, *sin_tab;
    int *topo_i;
};

// `regions`: reals taken by the wavefronts' own regions at the start of LDS.
template <typename T>
__device__ __forceinline__ WShared<T> wave_stage(const RArgs<T> &A, const Rob<T> &R0, unsigned char *raw, bool bwd,
                                 size_t regions) {
    WShared<T> sh;
    T *p = reinterpret_cast<T *>(raw) + regions;
    sh.wl = p; p += wave_wl4_reals(A.max_i);
    sh.topo_r = p; p += 3 * A.max_s;
    // sinusoid table over one period: phase 2 pi (t mod P) / P + 2 pi j / n_sin
    // evaluated in double precision (periodic in t, so equal to the direct form)
    const int n_tab = A.L.run_period * A.L.n_sin;
    sh.sin_tab = p; p += n_tab;
    sh.topo_i = reinterpret_cast<int *>(p);
    const int tid = threadIdx.x, BT = blockDim.x, S = R0.S, N = R0.N, O = R0.O;
    for (int q = tid; q < n_tab; q += BT) {
        const int tt = q / A.L.n_sin, j = q - tt * A.L.n_sin;
        sh.sin_tab[q] = T(::sin(2.0 * M_PI / A.L.run_period * (double)tt +
                                2.0 * M_PI / A.L.n_sin * (double)j));
    }
    for (int q = tid; q < S; q += BT) {
        sh.topo_r[q] = R0.L0[q]; sh.topo_r[S + q] = R0.k[q]; sh.topo_r[2 * S + q] = R0.coef[q];
        sh.topo_i[q] = R0.sa[q]; sh.topo_i[S + q] = R0.sb[q]; sh.topo_i[2 * S + q] = R0.sact[q];
    }
    const int v0 = R0.vptr[0], nv = R0.vptr[N] - v0;
    int *vp = sh.topo_i + 3 * S, *vx = vp + N + 1;
    for (int q = tid; q <= N; q += BT) vp[q] = R0.vptr[q] - v0;
    for (int q = tid; q < nv; q += BT) vx[q] = R0.vidx[v0 + q];
    const int a0 = R0.aptr[0], na = R0.aptr[O] - a0;
    int *ap = vx + nv, *ax = ap + O + 1;
    for (int q = tid; q <= O; q += BT) ap[q] = R0.aptr[q] - a0;
    for (int q = tid; q < na; q += BT) ax[q] = R0.aidx[a0 + q];
    {
        const int H = A.H, I = R0.I;
        for (int q = tid; q < I * H; q += BT) {
            const int i = q / H, j = q - i * H;
            sh.wl[wl4_at(i, j)] = R0.W1T[q];
        }
    }
    return sh;
}

template <typename T>
__device__ __forceinline__ void wave_point_at_shared(Rob<T> &R, const WShared<T> &sh) {
    const int S = R.S, N = R.N, O = R.O;
    const int nv = 2 * S;
    R.L0 = sh.topo_r; R.k = sh.topo_r + S; R.coef = sh.topo_r + 2 * S;
    R.sa = sh.topo_i; R.sb = sh.topo_i + S; R.sact = sh.topo_i + 2 * S;
    R.vptr = sh.topo_i + 3 * S;
    R.vidx = R.vptr + N + 1;
    R.aptr = R.vidx + nv;
    R.aidx = R.aptr + O + 1;
    R.WL = sh.wl;
}

template <typename T, int IMAX, int OMAX>
struct WRegs {
    T w1[IMAX];   // W1^T column of hidden unit `lane`
    T w2[OMAX];   // W2 column of hidden unit `lane`
    T b1;
};

template <typename T, int IMAX, int OMAX>
__device__ __forceinline__ void wave_load_weights(const Rob<T> &R, int H, int lane,
                                                  WRegs<T, IMAX, OMAX> &w) {
    const bool on = lane < H;
#pragma unroll
    for (int i = 0; i < IMAX; ++i) w.w1[i] = (on && i < R.I) ? R.W1T[(size_t)i * H + lane] : T(0);
#pragma unroll
    for (int o = 0; o < OMAX; ++o) w.w2[o] = (on && o < R.O) ? R.W2[(size_t)o * H + lane] : T(0);
    w.b1 = on ? R.b1[lane] : T(0);
}

// Per-lane copy of the wave's robot topology, held in registers for the
// whole rollout so the spring and vertex phases issue their state loads at
// once instead of first fetching indices from LDS: springs lane + 64 k
// (k < TSK) and the incident-spring entries of vertex `lane` (the first TDK;
// the wave kernel has N <= 64).  Springs and entries beyond those are read
// from the LDS copy.
constexpr int TSK = 6, TDK = 24;
#ifndef DL_FWD_SK
#define DL_FWD_SK 6
#endif
#ifndef DL_FWD_DK
#define DL_FWD_DK 24
#endif
#ifndef DL_BWD_SK
#define DL_BWD_SK 0
#endif
#ifndef DL_BWD_DK
#define DL_BWD_DK 0
#endif
constexpr int FSK = DL_FWD_SK, FDK = DL_FWD_DK;   // register depth of the forward kernel
constexpr int BSK = DL_BWD_SK, BDK = DL_BWD_DK;   // and of the adjoint kernel
template <typename T>
struct LaneTopo {
    int sp[TSK];                 // a | b << 8 | actuator << 16 (0 past the last spring)
    T L0[TSK], k[TSK], coef[TSK];
    int inc[TDK];                // vidx entries (spring << 1 | side), 0 past the degree
    int deg;                     // incident springs of vertex `lane` (0 when lane >= N)
};

template <typename T>
__device__ __forceinline__ void lane_topo_load(const Rob<T> &R, int lane, LaneTopo<T> &tp) {
#pragma unroll
    for (int q = 0; q < TSK; ++q) {
        const int s = lane + 64 * q;
        const bool on = s < R.S;
        tp.sp[q] = on ? (R.sa[s] | (R.sb[s] << 8) | (R.sact[s] << 16)) : 0;
        tp.L0[q] = on ? R.L0[s] : T(1);
        tp.k[q] = on ? R.k[s] : T(0);
        tp.coef[q] = on ? R.coef[s] : T(0);
    }
    const int e0 = lane < R.N ? R.vptr[lane] : 0;
    tp.deg = lane < R.N ? R.vptr[lane + 1] - e0 : 0;
#pragma unroll
    for (int q = 0; q < TDK; ++q) tp.inc[q] = q < tp.deg ? R.vidx[e0 + q] : 0;
}

// Rows of 4 reals at p + 4 r (16-byte aligned): components 0..2 of row r.
template <typename T>
__device__ __forceinline__ void row3_put(T *p, int r, T a, T b, T c) {
    p[4 * r] = a; p[4 * r + 1] = b; p[4 * r + 2] = c;
}
__device__ __forceinline__ void row3_put(float *p, int r, float a, float b, float c) {
    *reinterpret_cast<float4 *>(p + 4 * r) = make_float4(a, b, c, 0.f);
}
template <typename T>
__device__ __forceinline__ void row3_get(const T *p, int r, T &a, T &b, T &c) {
    a = p[4 * r]; b = p[4 * r + 1]; c = p[4 * r + 2];
}
__device__ __forceinline__ void row3_get(const float *p, int r, float &a, float &b, float &c) {
    const float4 v = *reinterpret_cast<const float4 *>(p + 4 * r);
    a = v.x; b = v.y; c = v.z;
}

// The wave kernels keep vertex i's position, velocity and adjoints as LDS
// rows of 4 at 4 i (one 16-byte access per vertex); the state, checkpoints
// and scratch rows in global memory are [3][N].
template <typename T>
__device__ __forceinline__ void rows_to_soa(T *g, const T *p, int N, int lane) {
    for (int i = lane; i < N; i += 64) {
        T a, b, c;
        row3_get(p, i, a, b, c);
        g[i] = a; g[N + i] = b; g[2 * N + i] = c;
    }
}
template <typename T>
__device__ __forceinline__ void soa_to_rows(T *p, const T *g, int N, int lane) {
    for (int i = lane; i < N; i += 64) row3_put(p, i, g[i], g[N + i], g[2 * N + i]);
}

// Spring impulses J (J = m.J, row s = spring s, [S][4]) of the current state.  Spring
// lane + 64 k from registers (loads issued unconditionally -- springs past
// the last read vertex 0 -- and stores masked), the rest from LDS.
template <int SK, typename T>
__device__ __forceinline__ void spring_impulses(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                                                const LaneTopo<T> &tp, int lane) {
    const int N = R.N, S = R.S;
    auto one = [&](int s, int a, int b, int ac, T L0, T kk, T co) {
        T xa[3], xb[3], va[3], vb[3];
        row3_get(m.xs, a, xa[0], xa[1], xa[2]);
        row3_get(m.xs, b, xb[0], xb[1], xb[2]);
        row3_get(m.vs, a, va[0], va[1], va[2]);
        row3_get(m.vs, b, vb[0], vb[1], vb[2]);
        const T d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
        const T wv[3] = {va[0] - vb[0], va[1] - vb[1], va[2] - vb[2]};
        T J[3];
        spring_impulse<T, true>(d, wv, L0, kk, co, m.outb[ac], A.P, J);
        if (s < S) row3_put(m.J, s, J[0], J[1], J[2]);
    };
#pragma unroll
    for (int q = 0; q < SK; ++q) {
        const int p = tp.sp[q];
        one(lane + 64 * q, p & 0xff, (p >> 8) & 0xff, p >> 16, tp.L0[q], tp.k[q], tp.coef[q]);
    }
#pragma unroll 2
    for (int s = lane + 64 * SK; s < S; s += 64)
        one(s, R.sa[s], R.sb[s], R.sact[s], R.L0[s], R.k[s], R.coef[s]);
}

// Impulse on vertex i (= lane): signed sum of J over its incident springs,
// the first TDK entries from registers (all loads at once, entries past the
// degree masked by selection), the rest from LDS.
template <int DK, typename T>
__device__ __forceinline__ void gather_J(const Rob<T> &R, const WSm<T> &m, const LaneTopo<T> &tp,
                                         int i, T vi[3]) {
    vi[0] = vi[1] = vi[2] = T(0);
#pragma unroll
    for (int q = 0; q < DK; ++q) {
        const int ent = tp.inc[q], s = ent >> 1;
        T j0, j1, j2;
        row3_get(m.J, s, j0, j1, j2);
        // entry sign, 0 past the degree: every load's value is used, so the
        // loads cannot be sunk into per-entry branches (products are exact)
        const T sg = q < tp.deg ? ((ent & 1) ? T(1) : T(-1)) : T(0);
        vi[0] += sg * j0;
        vi[1] += sg * j1;
        vi[2] += sg * j2;
    }
    // the rest of the list four entries at a time: the four indices, then
    // their rows, then the sums (groups pinned), so each round costs two
    // LDS round trips instead of one per index and one per row
    const int e1 = R.vptr[i + 1];
    int e = R.vptr[i] + DK;
    for (; e + 4 <= e1; e += 4) {
        int ent[4];
        T jr[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q) ent[q] = R.vidx[e + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) row3_get(m.J, ent[q] >> 1, jr[q][0], jr[q][1], jr[q][2]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const T sg = (ent[q] & 1) ? T(1) : T(-1);
            vi[0] += sg * jr[q][0]; vi[1] += sg * jr[q][1]; vi[2] += sg * jr[q][2];
        }
        __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 4, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 12, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 4, 0);
        __builtin_amdgcn_sched_group_barrier(kSchedValu, 24, 0);
    }
    for (; e < e1; ++e) {
        const int ent = R.vidx[e], s = ent >> 1;
        const T sg = (ent & 1) ? T(1) : T(-1);
        T j0, j1, j2;
        row3_get(m.J, s, j0, j1, j2);
        vi[0] += sg * j0; vi[1] += sg * j1; vi[2] += sg * j2;
    }
}

// The network input of step t (layout of taichi_env.py:149-183) into m.inb,
// assembled per block of slots (no per-slot division).
template <typename T>
__device__ __forceinline__ void wave_sensors(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                                             const T *sin_tab, int lane, int t, T c0, T c1, T c2,
                                             T u0, T u1, T u2, const T *tg) {
    // c: centre of mass, u: its velocity
    const int N = R.N, I = R.I;
    {
        const Layout &L = A.L;
        const int tmod = t % L.run_period;
        for (int j = lane; j < L.n_sin; j += 64) m.inb[j] = sin_tab[tmod * L.n_sin + j];
        if (L.vcom && lane < 3) m.inb[L.n_sin + lane] = lane == 0 ? u0 : (lane == 1 ? u1 : u2);
        const int base = L.vert_base(), vs = L.vstride();
        const T isc = T(1) / T(L.offset_scale);
        for (int i = lane; i < N; i += 64) {
            T *o = m.inb + base + i * vs;
            o[0] = (m.xs[4 * i] - c0) * isc;
            o[1] = (m.xs[4 * i + 1] - c1) * isc;
            o[2] = (m.xs[4 * i + 2] - c2) * isc;
            for (int d = 3; d < vs; ++d) o[d] = T(0);
        }
        const int tail = base + N * vs, nv = L.dup_v * L.n_tv;
        for (int q = lane; q < I - tail; q += 64) {
            T val;
            if (q < nv) {
                const int mm = q % L.n_tv;
                val = tg[mm == 0 ? L.tv0 : (mm == 1 ? L.tv1 : L.tv2)] / T(L.max_speed);
            } else {
                val = q < nv + L.dup_h ? tg[3] : tg[4];
            }
            m.inb[tail + q] = val;
        }
    }
}

// one forward step of the wave's robot (MODE_FWD / MODE_REC like fwd_step).
// ROWS: a forward step also writes the step's adjoint rows (as MODE_REC does)
// at scratch slot `slot`, for the adjoint that reads them instead of
// recomputing.
template <typename T, int OMAX, int SK, int DK, bool ROWS = false>
__device__ __forceinline__ void wave_fwd_step(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                              const WRegs<T, 1, OMAX> &w, const LaneTopo<T> &tp,
                              const T *sin_tab, int lane, int t, int mode, int slot, T &lsum) {
    const int N = R.N, S = R.S, I = R.I, O = R.O, H = A.H;
    T s6[6] = {0, 0, 0, 0, 0, 0};
    const T invN = T(1) / T(N);
    for (int i = lane; i < N; i += 64) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            s6[d] += m.xs[4 * i + d] * invN;
            s6[3 + d] += m.vs[4 * i + d] * invN;
        }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) s6[q] = wave_sum(s6[q]);
    const T *tg = tgt_at(R, t);
    if (mode == MODE_FWD) {
        if (t % A.K == 0) {
            T *ck = R.ckpt + (size_t)(t / A.K) * 6 * N;
            rows_to_soa(ck, m.xs, N, lane);
            rows_to_soa(ck + 3 * N, m.vs, N, lane);
        }
        if (A.store_traj && R.traj) {
            T *row = R.traj + (size_t)t * (3 * N + 3);
            rows_to_soa(row, m.xs, N, lane);
            if (lane < 3) row[3 * N + lane] = lane == 0 ? s6[0] : (lane == 1 ? s6[1] : s6[2]);
        }
        if (A.loss_kind == 1 && t > R.Th - A.window) lsum += task_term(s6, tg);
    }
    const bool rows = ROWS || mode == MODE_REC;
    if (rows) {
        T *sx = R.sx + (size_t)slot * 6 * N;
        rows_to_soa(sx, m.xs, N, lane);
        rows_to_soa(sx + 3 * N, m.vs, N, lane);
        if (lane < 6) {
            T val = s6[0];
#pragma unroll
            for (int q = 1; q < 6; ++q) val = lane == q ? s6[q] : val;
            R.scv[slot * 6 + lane] = val;
        }
    }
    wave_sensors(A, R, m, sin_tab, lane, t, s6[0], s6[1], s6[2], s6[3], s6[4], s6[5], tg);
    wave_sync();
    if (rows)
        for (int q = lane; q < I; q += 64) R.sin[(size_t)slot * I + q] = m.inb[q];
    // layer 1: lane j = hidden unit j, W1 column j from the LDS copy (H = HMAX
    // in the wave kernels), inputs broadcast from LDS, input i into a[i % 4].
    // The forward and the segment recompute run this same code, so the
    // recompute reproduces the forward's hidden values bit for bit.
    T acc = w.b1;
    if (R.WL) {
        // 16 inputs per round: one 16-byte read of W1 per lane and group of
        // four inputs (the LDS copy's grouped layout), the inputs as 16-byte
        // broadcasts (m.inb is 16-byte aligned); input i goes into a[i % 4].
        // The scheduling groups keep a round's 8 reads ahead of its 16 FMAs
        // (left to itself the compiler reuses registers and waits on every
        // read once the kernel's register pressure grows).
        static_assert(std::is_same<T, float>::value, "LDS copy of W1: FP32 wave kernels");
        T a[4] = {acc, T(0), T(0), T(0)};
        const float4 *w4 = reinterpret_cast<const float4 *>(R.WL) + lane;
        const float4 *x4 = reinterpret_cast<const float4 *>(m.inb);
        int i = 0;
        for (; i + 16 <= I; i += 16) {
            float4 wv[4], xv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) wv[q] = w4[((i >> 2) + q) * (HMAX + 1)];
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = x4[(i >> 2) + q];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[0] += wv[q].x * xv[q].x;
                a[1] += wv[q].y * xv[q].y;
                a[2] += wv[q].z * xv[q].z;
                a[3] += wv[q].w * xv[q].w;
            }
            __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 8, 0);
            __builtin_amdgcn_sched_group_barrier(kSchedValu, 16, 0);
        }
        for (; i < I; ++i) a[i & 3] += R.WL[wl4_at(i, lane)] * m.inb[i];
        acc = (a[0] + a[1]) + (a[2] + a[3]);
    } else if (lane < H) {
#pragma unroll 8
        for (int i = 0; i < I; ++i) acc += R.W1T[(size_t)i * H + lane] * m.inb[i];
    }
    T hj, dhj;
    activate<T>(A.act, acc, hj, dhj);
    if (lane >= H) hj = dhj = T(0);
    if (rows && lane < H) {
        R.shid[(size_t)slot * H + lane] = hj;
        R.sdhid[(size_t)slot * H + lane] = dhj;
    }
    // layer 2: one wave reduction per output, lane o then activates output o
    {
        T zsel = T(0);
#pragma unroll
        for (int o = 0; o < OMAX; ++o) {
            if (o < O) {
                const T z = wave_sum(w.w2[o] * hj) + R.b2[o];
                zsel = lane == o ? z : zsel;
            }
        }
        T y, dy;
        activate<T>(A.act, zsel, y, dy);
        if (lane < O) {
            m.outb[lane] = y;
            m.dout[lane] = dy;
            if (rows) {
                R.sout[(size_t)slot * O + lane] = y;
                R.sdout[(size_t)slot * O + lane] = dy;
            }
        }
    }
    wave_sync();
    spring_impulses<SK>(A, R, m, tp, lane);
    wave_sync();
    for (int i = lane; i < N; i += 64) {
        T vi[3];
        gather_J<DK>(R, m, tp, i, vi);
        if constexpr (ROWS) {
            // kept for the adjoint, which then skips the spring pass
            T *sv = R.svi + (size_t)slot * 3 * N;
            sv[i] = vi[0]; sv[N + i] = vi[1]; sv[2 * N + i] = vi[2];
        }
        T x[3], v[3];
        row3_get(m.xs, i, x[0], x[1], x[2]);
        row3_get(m.vs, i, v[0], v[1], v[2]);
        T xn[3], vn[3];
        integrate<T>(x, v, vi, A.P, xn, vn);
        row3_put(m.xs, i, xn[0], xn[1], xn[2]);
        row3_put(m.vs, i, vn[0], vn[1], vn[2]);
    }
    wave_sync();
}

// ROWS: also write every step's adjoint rows into the robot's scratch (slot =
// step; the scratch holds K >= horizon rows).
template <typename T, int OMAX, bool ROWS>
__global__ void __launch_bounds__(256) k_wave_fwd(RArgs<T> A, int n_robots) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int r0 = blockIdx.x * (blockDim.x >> 6);
    const Rob<T> R0 = make_rob(A, r0);
    const WShared<T> sh = wave_stage(A, R0, smem_raw, false,
                                     (blockDim.x >> 6) * wave_region(A.max_n, A.max_s, A.max_i, A.max_o, false));
    __syncthreads();
    const int r = r0 + wid;
    if (r >= n_robots) return;
    Rob<T> R = make_rob(A, r);
    WSm<T> m = wave_carve(A, reinterpret_cast<T *>(smem_raw), wid, false);
    if (A.robots[r].spring0 != A.robots[r0].spring0 || A.robots[r].weights != A.robots[r0].weights) {
        // robots of one workgroup must share a design (the host packs them so)
        if (lane == 0 && A.loss) A.loss[r] = T(NAN);
        if (lane == 0 && A.err) atomicOr(A.err, DL_ROLLOUT_ERR_PACKING);
        return;
    }
    wave_point_at_shared(R, sh);
    for (int q = lane; q < (A.max_i > 192 ? A.max_i : 192) + 4; q += 64) m.inb[q] = T(0);
    const int N = R.N;
    const int t0 = A.robots[r].t_fwd;
    int t1 = R.Th;
    if (A.max_steps > 0 && t0 + A.max_steps < t1) t1 = t0 + A.max_steps;
    if (t0 >= t1) return;
    if (ROWS && (R.Th > A.K || R.sx == nullptr)) {
        // the stored rows need K >= horizon rows of scratch
        if (lane == 0 && A.loss) A.loss[r] = T(NAN);
        if (lane == 0 && A.err) atomicOr(A.err, DL_ROLLOUT_ERR_HORIZON);
        return;
    }
    WRegs<T, 1, OMAX> w;
    wave_load_weights(R, A.H, lane, w);
    LaneTopo<T> tp;
    lane_topo_load(R, lane, tp);
    soa_to_rows(m.xs, R.state, N, lane);
    soa_to_rows(m.vs, R.state + 3 * N, N, lane);
    wave_sync();
    T lsum = T(0);
    for (int t = t0; t < t1; ++t)
        wave_fwd_step<T, OMAX, FSK, FDK, ROWS>(A, R, m, w, tp, sh.sin_tab, lane, t, MODE_FWD, t, lsum);
    if (t1 == R.Th) {
        T s6[6] = {0, 0, 0, 0, 0, 0};
        const T invN = T(1) / T(N);
        for (int i = lane; i < N; i += 64)
            for (int d = 0; d < 3; ++d) {
                s6[d] += m.xs[4 * i + d] * invN;
                s6[3 + d] += m.vs[4 * i + d] * invN;
            }
        for (int q = 0; q < 6; ++q) s6[q] = wave_sum(s6[q]);
        if (A.loss_kind == 1) lsum += task_term(s6, tgt_at(R, R.Th));
        if (A.store_traj && R.traj) {
            T *row = R.traj + (size_t)R.Th * (3 * N + 3);
            rows_to_soa(row, m.xs, N, lane);
        }
    }
    rows_to_soa(R.state, m.xs, N, lane);
    rows_to_soa(R.state + 3 * N, m.vs, N, lane);
    if (A.loss_kind == 1 && lane == 0) A.loss[r] += lsum / T(A.window);
}

// Per-lane copy of one step's recomputed intermediates, loaded one step ahead
// of its use so the global-memory latency overlaps the previous adjoint step.
// Holds 6N <= 6*64 state values (the wave kernel is limited to N <= 64).
// Each lane reads back exactly the elements it stored during the recompute.
template <typename T>
struct Pref {
    T x[3], v[3];
    T vi[3];            // impulse sums (stored rows only)
    T out, dout, dh, cv;
};

template <bool VI = false, typename T>
__device__ __forceinline__ void pref_load(const Rob<T> &R, int H, int lane, int slot, Pref<T> &p) {
    const int N = R.N, O = R.O;
    const T *sx = R.sx + (size_t)slot * 6 * N;
    const int i = lane < N ? lane : 0;   // lane = vertex (N <= 64)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        p.x[d] = sx[d * N + i];
        p.v[d] = sx[3 * N + d * N + i];
    }
    if constexpr (VI) {
        const T *sv = R.svi + (size_t)slot * 3 * N;
#pragma unroll
        for (int d = 0; d < 3; ++d) p.vi[d] = sv[d * N + i];
    }
    p.out = lane < O ? R.sout[(size_t)slot * O + lane] : T(0);
    p.dout = lane < O ? R.sdout[(size_t)slot * O + lane] : T(0);
    p.dh = lane < H ? R.sdhid[(size_t)slot * H + lane] : T(0);
    p.cv = lane < 6 ? R.scv[slot * 6 + lane] : T(0);
}

// acc[p] += rows[p][4 (j0 + k)] * z_k in order k = 0..15 (rows[p] points at
// W1[0][i] of the lane's input i in the grouped LDS copy, where W1[j][i] is
// 4 j further), z_k from lane k of the lane's 16-lane row of z.  The 16 NP
// entries are read from LDS into registers before the FMAs (order pinned by
// scheduling groups).
template <int NP, int... K>
__device__ __forceinline__ void in_adj_rows(float (&acc)[NP], const float *const (&rows)[NP], int j0,
                                            float z, std::integer_sequence<int, K...>) {
    float rv[NP][16];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int k = 0; k < 16; ++k) rv[p][k] = rows[p][4 * (j0 + k)];
    auto one = [&](float zk, int k) {
#pragma unroll
        for (int p = 0; p < NP; ++p) acc[p] += rv[p][k] * zk;
    };
    (one(row_bcast<K>(z), K), ...);
    __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 16 * NP, 0);
    __builtin_amdgcn_sched_group_barrier(kSchedValu, 16 * (NP + 1), 0);
}

// VI: the step's impulse sums come from the rows the forward kept (no spring
// pass and gather here).
template <typename T, int IMAX, int OMAX, int SK, int DK, bool VI = false>
__device__ __forceinline__ void wave_bwd_step(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m, int r, int lane,
                              int t, int slot, Pref<T> &pf, const WRegs<T, 1, OMAX> &w,
                              const LaneTopo<T> &tp, int next_slot) {
    const int N = R.N, S = R.S, I = R.I, O = R.O, H = A.H;
    // consume the prefetched step (lane = vertex)
    if (lane < N) {
        row3_put(m.xs, lane, pf.x[0], pf.x[1], pf.x[2]);
        row3_put(m.vs, lane, pf.v[0], pf.v[1], pf.v[2]);
    }
    if (lane < O) {
        m.outb[lane] = pf.out;
        m.dout[lane] = pf.dout;
    }
    T cvr[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) cvr[q] = lane_read(pf.cv, q);
    const T dhj = pf.dh, pf_dout = pf.dout;
    const T pvi[3] = {pf.vi[0], pf.vi[1], pf.vi[2]};
    if (next_slot >= 0) pref_load<VI>(R, H, lane, next_slot, pf);   // in flight during this step
    wave_sync();
    if constexpr (!VI) {
        spring_impulses<SK>(A, R, m, tp, lane);
        wave_sync();
    }
    for (int i = lane; i < N; i += 64) {
        T vi[3];
        if constexpr (VI) {
            vi[0] = pvi[0]; vi[1] = pvi[1]; vi[2] = pvi[2];   // lane = vertex (N <= 64)
        } else {
            gather_J<DK>(R, m, tp, i, vi);
        }
        T x[3], v[3];
        row3_get(m.xs, i, x[0], x[1], x[2]);
        row3_get(m.vs, i, v[0], v[1], v[2]);
        T xnb[3], vnb[3];
        row3_get(m.xb, i, xnb[0], xnb[1], xnb[2]);
        row3_get(m.vb, i, vnb[0], vnb[1], vnb[2]);
        T xb[3], vb[3], vib[3];
        integrate_adj<T>(x, v, vi, A.P, xnb, vnb, xb, vb, vib);
        row3_put(m.xb, i, xb[0], xb[1], xb[2]);
        row3_put(m.vb, i, vb[0], vb[1], vb[2]);
        row3_put(m.vib, i, vib[0], vib[1], vib[2]);
    }
    wave_sync();
    // spring adjoints G, rows of 4: row 2 s the length-vector adjoint of
    // spring s, row 2 s + 1 its relative-velocity adjoint; the actuation
    // adjoints are summed per actuator in registers (lane partial sums, then
    // one wave sum per actuator)
    T *G = m.J;
    T abp[OMAX];
#pragma unroll
    for (int o = 0; o < OMAX; ++o) abp[o] = T(0);
    auto adj_one = [&](int s, int a, int b, int ac, T L0, T kk, T co) {
        T xa[3], xq[3], va[3], vq[3], ja[3], jq[3];
        row3_get(m.xs, a, xa[0], xa[1], xa[2]);
        row3_get(m.xs, b, xq[0], xq[1], xq[2]);
        row3_get(m.vs, a, va[0], va[1], va[2]);
        row3_get(m.vs, b, vq[0], vq[1], vq[2]);
        row3_get(m.vib, a, ja[0], ja[1], ja[2]);
        row3_get(m.vib, b, jq[0], jq[1], jq[2]);
        const T d[3] = {xa[0] - xq[0], xa[1] - xq[1], xa[2] - xq[2]};
        const T wv[3] = {va[0] - vq[0], va[1] - vq[1], va[2] - vq[2]};
        const T Jb[3] = {jq[0] - ja[0], jq[1] - ja[1], jq[2] - ja[2]};
        T gd[3], gw[3], ga;
        spring_impulse_adj<T, true>(d, wv, L0, kk, co, m.outb[ac], A.P, Jb, gd, gw, ga);
        if (s < S) {
            row3_put(G, 2 * s, gd[0], gd[1], gd[2]);
            row3_put(G, 2 * s + 1, gw[0], gw[1], gw[2]);
#pragma unroll
            for (int o = 0; o < OMAX; ++o) abp[o] += ac == o ? ga : T(0);
        }
    };
#pragma unroll
    for (int q = 0; q < SK; ++q) {
        const int p = tp.sp[q];
        adj_one(lane + 64 * q, p & 0xff, (p >> 8) & 0xff, p >> 16, tp.L0[q], tp.k[q], tp.coef[q]);
    }
#pragma unroll 2
    for (int s = lane + 64 * SK; s < S; s += 64)
        adj_one(s, R.sa[s], R.sb[s], R.sact[s], R.L0[s], R.k[s], R.coef[s]);
    wave_sync();
    for (int i = lane; i < N; i += 64) {
        T gx[3] = {0, 0, 0}, gv[3] = {0, 0, 0};
#pragma unroll
        for (int q = 0; q < DK; ++q) {
            const int ent = tp.inc[q], s = ent >> 1;
            T g[6];
            row3_get(G, 2 * s, g[0], g[1], g[2]);
            row3_get(G, 2 * s + 1, g[3], g[4], g[5]);
            // entry sign, 0 past the degree (loads stay unconditional)
            const T sg = q < tp.deg ? ((ent & 1) ? T(-1) : T(1)) : T(0);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gx[c] += sg * g[c];
                gv[c] += sg * g[3 + c];
            }
        }
        // four entries per round (indices, then rows, then sums)
        const int e1 = R.vptr[i + 1];
        int e = R.vptr[i] + DK;
        for (; e + 4 <= e1; e += 4) {
            int ent[4];
            T gr[4][6];
#pragma unroll
            for (int q = 0; q < 4; ++q) ent[q] = R.vidx[e + q];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int s = ent[q] >> 1;
                row3_get(G, 2 * s, gr[q][0], gr[q][1], gr[q][2]);
                row3_get(G, 2 * s + 1, gr[q][3], gr[q][4], gr[q][5]);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const T sg = (ent[q] & 1) ? T(-1) : T(1);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    gx[c] += sg * gr[q][c];
                    gv[c] += sg * gr[q][3 + c];
                }
            }
            __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 4, 0);
            __builtin_amdgcn_sched_group_barrier(kSchedValu, 16, 0);
            __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 8, 0);
            __builtin_amdgcn_sched_group_barrier(kSchedValu, 40, 0);
        }
        for (; e < e1; ++e) {
            const int ent = R.vidx[e], s = ent >> 1;
            const T sg = (ent & 1) ? T(-1) : T(1);
            T g[6];
            row3_get(G, 2 * s, g[0], g[1], g[2]);
            row3_get(G, 2 * s + 1, g[3], g[4], g[5]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                gx[c] += sg * g[c];
                gv[c] += sg * g[3 + c];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xb[4 * i + d] += gx[d];
            m.vb[4 * i + d] += gv[d];
        }
    }
#pragma unroll
    for (int o = 0; o < OMAX; ++o) {
        if (o < O) {
            const T z = wave_sum(abp[o]) * lane_read(pf_dout, o);
            if (lane == o) {
                m.z2b[o] = z;
                R.sz2[(size_t)slot * O + o] = z;
            }
        }
    }
    wave_sync();
    T z1;
    {
        T hb = T(0);
#pragma unroll
        for (int o = 0; o < OMAX; ++o)
            if (o < O) hb += w.w2[o] * m.z2b[o];
        z1 = hb * dhj;
        if (lane < H) R.sz1[(size_t)slot * H + lane] = z1;
    }
    // input adjoint inbar = W1^T z1: lane i + 64 p sums row i of the LDS copy
    // against z1, all row blocks in one pass over j, in chunks of 16: each
    // chunk reads z1[j0 .. j0 + 15] into every 16-lane row with one LDS read
    // and takes each z1_j from it by a row broadcast.  A broadcast reads 0
    // from a disabled lane, so every lane runs the loop (rows past I read row
    // I-1, stores guarded).  The wave kernels run with H = HMAX (checked at
    // launch).
    if (std::is_same<T, float>::value && R.WL) {
        constexpr int NP = (IMAX + 63) / 64;
        m.z1b[lane] = z1;
        wave_sync();
        const float *rows[NP];
        float acc[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const int i = lane + 64 * p;
            rows[p] = reinterpret_cast<const float *>(R.WL) + wl4_at(i < I ? i : I - 1, 0);
            acc[p] = 0.f;
        }
#pragma unroll 1
        for (int j0 = 0; j0 < HMAX; j0 += 16) {
            const float zc = m.z1b[j0 + (lane & 15)];
            in_adj_rows(acc, rows, j0, zc, std::make_integer_sequence<int, 16>{});
        }
#pragma unroll
        for (int p = 0; p < NP; ++p)
            if (lane + 64 * p < I) m.inbar[lane + 64 * p] = T(acc[p]);
    } else {
        m.z1b[lane] = z1;
        wave_sync();
        for (int i = lane; i < I; i += 64) {
            T acc = T(0);
#pragma unroll 16
            for (int j = 0; j < H; ++j) acc += R.W1[(size_t)j * I + i] * m.z1b[j];
            m.inbar[i] = acc;
        }
    }
    wave_sync();
    const Layout &L = A.L;
    const int base = L.vert_base(), vs = L.vstride();
    const T inv = T(1) / T(L.offset_scale);
    T cb[3] = {0, 0, 0};
    const T *eg = R.trajg ? R.trajg + (size_t)t * (3 * N + 3) : nullptr;
    for (int i = lane; i < N; i += 64) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T g = m.inbar[base + i * vs + d] * inv;
            m.xb[4 * i + d] += g + (eg ? eg[d * N + i] : T(0));
            cb[d] -= g;
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) cb[d] = wave_sum(cb[d]);
    T vcb[3] = {0, 0, 0};
    if (L.vcom)
        for (int d = 0; d < 3; ++d) vcb[d] = m.inbar[L.n_sin + d];
    if (eg)
        for (int d = 0; d < 3; ++d) cb[d] += eg[3 * N + d];
    if (A.loss_kind == 1 && t > R.Th - A.window) {
        const T lg = A.loss_grad[r] / T(A.window);
        const T *tg = tgt_at(R, t);
        vcb[0] += T(2) * lg * (cvr[3] - tg[0]);
        vcb[2] += T(2) * lg * (cvr[5] - tg[2]);
        cb[1] += T(2) * lg * (cvr[1] - tg[3]);
    }
    const T invN = T(1) / T(N);
    for (int i = lane; i < N; i += 64) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            m.xb[4 * i + d] += cb[d] * invN;
            m.vb[4 * i + d] += vcb[d] * invN;
        }
    }
    wave_sync();
}

// Weight-gradient fold of n steps starting at scratch slot `off`:
// W1g[j][i] += sum_s z1[s][j] in[s][i], then the b1, W2 and b2 gradients.
// Lanes run over inputs i with 64 accumulators (one per hidden unit).  The
// steps go in chunks: a chunk's rows (z1 and the lane's inputs) are loaded
// with every load issued before any is used, staged in LDS (`zl`, zcap
// reals that are dead while the fold runs), then summed from LDS (z1 as
// broadcasts).  The rows were written a segment earlier and come from HBM,
// so one load in flight at a time would cost a memory round trip per step.
// The loads walk per-lane pointers (the hidden rows are HMAX wide in the
// wave kernels, so their offsets are immediates): the kernel has no scalar
// registers to spare for a row address per load.  Steps are summed in
// ascending order.  Only the W1 rows of inputs [ia, ib) are folded, and the
// b1, W2 and b2 gradients only with `part2`, so two wavefronts can fold one
// segment between them, each into entries of its own.
constexpr int kFoldChunk = 16;
// LDS reals per staged step: z1 and input rows (first pass); z1, hidden and
// z2 rows (second pass, z2 at most 8 wide)
constexpr int kFoldStepReals = 2 * HMAX + 8;

// Reals of a wave region free while the fold runs: its spring rows and its
// state rows (`bwd`: the adjoint region, else the recompute region), which
// lie back to back at the start of the region.
__host__ __device__ inline int wave_fold_cap(int max_n, int max_s, bool bwd) {
    return (bwd ? 8 : 4) * max_s + 8 * max_n;
}

template <int OMAX, typename T>
__device__ __forceinline__ void wave_seg_wgrad(const RArgs<T> &A, const Rob<T> &R, T *zl, int zcap,
                                               int lane, int off, int n, int ia = 0,
                                               int ib = 1 << 30, bool part2 = true) {
    static_assert(OMAX <= kFoldStepReals - 2 * HMAX, "z2 rows wider than the fold's staging");
    const int I = R.I, O = R.O;
    constexpr int CB = kFoldChunk;
    const int chunk = min(zcap / (2 * HMAX), CB);
    T *const xl = zl + chunk * HMAX;
    for (int i0 = ia; i0 < min(I, ib); i0 += 64) {
        const int i = i0 + lane;
        const bool ok = i < I;
        T acc[HMAX];
#pragma unroll
        for (int j = 0; j < HMAX; ++j) acc[j] = T(0);
        const T *zp = R.sz1 + (size_t)off * HMAX + lane;
        const T *xp = R.sin + (size_t)off * I + (ok ? i : 0);
        for (int c0 = 0; c0 < n; c0 += chunk) {
            const int cn = min(chunk, n - c0);
            T zr[CB], xr[CB];
            const T *xq = xp;
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                zr[u] = xr[u] = T(0);
                if (u < cn) {
                    zr[u] = zp[u * HMAX];
                    xr[u] = *xq;
                }
                xq += I;
            }
            zp += cn * HMAX;
            xp += (size_t)cn * I;
            wave_sync();   // the previous chunk's reads of the stage are done
#pragma unroll
            for (int u = 0; u < CB; ++u) {
                if (u < cn) {
                    zl[u * HMAX + lane] = zr[u];
                    xl[u * HMAX + lane] = xr[u];
                }
            }
            wave_sync();
            for (int u = 0; u < cn; ++u) {
                const T x = xl[u * HMAX + lane];
                const T *z = zl + u * HMAX;
                if constexpr (std::is_same<T, float>::value) {
                    // z1 row as 16-byte broadcasts, 16 entries per pinned
                    // round (64 reads in flight beside the 64 accumulators
                    // would not fit; scalar reads each wait on the last)
#pragma unroll
                    for (int j0 = 0; j0 < HMAX; j0 += 16) {
                        float4 zz[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            zz[q] = *reinterpret_cast<const float4 *>(z + j0 + 4 * q);
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            acc[j0 + 4 * q + 0] += zz[q].x * x;
                            acc[j0 + 4 * q + 1] += zz[q].y * x;
                            acc[j0 + 4 * q + 2] += zz[q].z * x;
                            acc[j0 + 4 * q + 3] += zz[q].w * x;
                        }
                        __builtin_amdgcn_sched_group_barrier(kSchedDsRead, 4, 0);
                        __builtin_amdgcn_sched_group_barrier(kSchedValu, 16, 0);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < HMAX; ++j) acc[j] += z[j] * x;
                }
            }
        }
        if (ok) {
            // 16 entries' loads in flight, then their stores (entry by entry,
            // each load would wait behind the store before it: rows j I + i
            // may alias for all the compiler knows)
            T *const wp = R.Wg1 + i;
#pragma unroll
            for (int j0 = 0; j0 < HMAX; j0 += 16) {
                T g[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) g[q] = wp[(size_t)(j0 + q) * I];
#pragma unroll
                for (int q = 0; q < 16; ++q) wp[(size_t)(j0 + q) * I] = g[q] + acc[j0 + q];
            }
        }
    }
    if (!part2) return;
    // lane j sums z1[s][j] and z2[s][o] hid[s][j]; lane o < O sums z2[s][o]
    const int chunk2 = min(zcap / kFoldStepReals, CB);
    T *const hl = zl + chunk2 * HMAX, *const ql = hl + chunk2 * HMAX;
    T b = T(0), c = T(0), a[OMAX];
#pragma unroll
    for (int o = 0; o < OMAX; ++o) a[o] = T(0);
    const T *zp = R.sz1 + (size_t)off * HMAX + lane;
    const T *hp = R.shid + (size_t)off * HMAX + lane;
    const T *qp = R.sz2 + (size_t)off * O + (lane < O ? lane : 0);
    for (int c0 = 0; c0 < n; c0 += chunk2) {
        const int cn = min(chunk2, n - c0);
        T zr[CB], hr[CB], qr[CB];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            zr[u] = hr[u] = qr[u] = T(0);
            if (u < cn) {
                zr[u] = zp[u * HMAX];
                hr[u] = hp[u * HMAX];
                qr[u] = qp[u * O];
            }
        }
        zp += cn * HMAX;
        hp += cn * HMAX;
        qp += cn * O;
        wave_sync();
#pragma unroll
        for (int u = 0; u < CB; ++u) {
            if (u < cn) {
                zl[u * HMAX + lane] = zr[u];
                hl[u * HMAX + lane] = hr[u];
                if (lane < OMAX) ql[u * OMAX + lane] = lane < O ? qr[u] : T(0);
            }
        }
        wave_sync();
        for (int u = 0; u < cn; ++u) {
            const T h = hl[u * HMAX + lane];
            const T *q = ql + u * OMAX;
            b += zl[u * HMAX + lane];
            c += q[lane < OMAX ? lane : 0];
#pragma unroll
            for (int o = 0; o < OMAX; ++o)
                if (o < O) a[o] += q[o] * h;
        }
    }
    R.bg1[lane] += b;
#pragma unroll
    for (int o = 0; o < OMAX; ++o)
        if (o < O) R.Wg2[(size_t)o * HMAX + lane] += a[o];
    if (lane < O) R.bg2[lane] += c;
}

// Adjoint state of the robot's last state (t_bwd) into LDS, with the loss
// terms of the final state when t_bwd is the horizon.
template <typename T>
__device__ __forceinline__ void wave_bwd_begin(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m, int r,
                                               int lane) {
    const int N = R.N;
    soa_to_rows(m.xb, R.adj, N, lane);
    soa_to_rows(m.vb, R.adj + 3 * N, N, lane);
    wave_sync();
    if (A.robots[r].t_bwd == R.Th) {
        T s6[6] = {0, 0, 0, 0, 0, 0};
        const T invN = T(1) / T(N);
        for (int i = lane; i < N; i += 64)
            for (int d = 0; d < 3; ++d) {
                s6[d] += R.state[d * N + i] * invN;
                s6[3 + d] += R.state[3 * N + d * N + i] * invN;
            }
        for (int q = 0; q < 6; ++q) s6[q] = wave_sum(s6[q]);
        T cb[3] = {0, 0, 0}, vcb[3] = {0, 0, 0};
        if (A.loss_kind == 1) {
            const T lg = A.loss_grad[r] / T(A.window);
            const T *tg = tgt_at(R, R.Th);
            vcb[0] = T(2) * lg * (s6[3] - tg[0]);
            vcb[2] = T(2) * lg * (s6[5] - tg[2]);
            cb[1] = T(2) * lg * (s6[1] - tg[3]);
        }
        const T *eg = R.trajg ? R.trajg + (size_t)R.Th * (3 * N + 3) : nullptr;
        for (int i = lane; i < N; i += 64)
            for (int d = 0; d < 3; ++d) {
                m.xb[4 * i + d] += cb[d] * invN + (eg ? eg[d * N + i] : T(0));
                m.vb[4 * i + d] += vcb[d] * invN;
            }
        wave_sync();
    }
}

// Recompute steps [s0, t1) of the segment starting at checkpoint s0 / K into
// the scratch rows of R (slot t - s0); ROWS: with the per-vertex impulse sums.
template <typename T, int OMAX, bool ROWS = false>
__device__ __forceinline__ void wave_recompute(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                                               const WRegs<T, 1, OMAX> &w, const LaneTopo<T> &tp,
                                               const T *sin_tab, int lane, int s0, int t1) {
    const int N = R.N;
    const T *ck = R.ckpt + (size_t)(s0 / A.K) * 6 * N;
    soa_to_rows(m.xs, ck, N, lane);
    soa_to_rows(m.vs, ck + 3 * N, N, lane);
    wave_sync();
    T dummy = T(0);
    for (int t = s0; t < t1; ++t)
        wave_fwd_step<T, OMAX, BSK, BDK, ROWS>(A, R, m, w, tp, sin_tab, lane, t, MODE_REC,
                                               t - s0, dummy);
    // the scratch stores complete before any lane reads them back
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
}

// Adjoint steps tb - 1 down to lo of the segment starting at s0 (rows of R
// recomputed); the steps' weight-gradient rows are left in R's scratch.
template <typename T, int IMAX, int OMAX, bool VI = false>
__device__ __forceinline__ void wave_bwd_steps(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                                               const WRegs<T, 1, OMAX> &w, const LaneTopo<T> &tp,
                                               int r, int lane, int s0, int lo, int tb) {
    Pref<T> pf;
    pref_load<VI>(R, A.H, lane, tb - 1 - s0, pf);
    for (int t = tb - 1; t >= lo; --t)
        wave_bwd_step<T, IMAX, OMAX, BSK, BDK, VI>(A, R, m, r, lane, t, t - s0, pf, w, tp,
                                                   t - 1 >= lo ? t - 1 - s0 : -1);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    wave_sync();
}

// The same, then the segment's weight-gradient fold (zl: LDS scratch of
// zcap >= HMAX reals).
template <typename T, int IMAX, int OMAX>
__device__ __forceinline__ void wave_bwd_segment(const RArgs<T> &A, const Rob<T> &R, WSm<T> &m,
                                                 const WRegs<T, 1, OMAX> &w, const LaneTopo<T> &tp,
                                                 int r, int lane, int s0, int lo, int tb) {
    wave_bwd_steps<T, IMAX, OMAX>(A, R, m, w, tp, r, lane, s0, lo, tb);
    wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, true), lane, lo - s0,
                         tb - lo);
    wave_sync();
}

template <typename T>
__device__ __forceinline__ void wave_bwd_end(const Rob<T> &R, const WSm<T> &m, int lane) {
    const int N = R.N;
    rows_to_soa(R.adj, m.xb, N, lane);
    rows_to_soa(R.adj + 3 * N, m.vb, N, lane);
}

// The whole adjoint of robot r by one wavefront: segments in turn, each
// recomputed from its checkpoint and then run in reverse.
template <typename T, int IMAX, int OMAX>
__device__ void wave_bwd_robot(const RArgs<T> &A, Rob<T> &R, WSm<T> &m, const WShared<T> &sh, int r,
                               int lane) {
    for (int q = lane; q < (A.max_i > 192 ? A.max_i : 192) + 4; q += 64) m.inb[q] = T(0);
    const int K = A.K;
    int tb = A.robots[r].t_bwd;
    if (tb <= 0) return;
    WRegs<T, 1, OMAX> w;
    wave_load_weights(R, A.H, lane, w);
    LaneTopo<T> tp;
    lane_topo_load(R, lane, tp);
    wave_bwd_begin(A, R, m, r, lane);
    int budget = A.max_steps > 0 ? A.max_steps : 0x7fffffff;
    while (tb > 0 && budget > 0) {
        const int s0 = (tb - 1) / K * K;
        const int lo = max(s0, tb - budget);
        wave_recompute(A, R, m, w, tp, sh.sin_tab, lane, s0, tb);
        wave_bwd_segment<T, IMAX, OMAX>(A, R, m, w, tp, r, lane, s0, lo, tb);
        budget -= tb - lo;
        tb = lo;
    }
    wave_bwd_end(R, m, lane);
}

template <typename T, int IMAX, int OMAX>
__global__ void __launch_bounds__(256) k_wave_bwd(RArgs<T> A, int n_robots) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int r0 = blockIdx.x * (blockDim.x >> 6);
    const Rob<T> R0 = make_rob(A, r0);
    const WShared<T> sh = wave_stage(A, R0, smem_raw, true,
                                     (blockDim.x >> 6) * wave_region(A.max_n, A.max_s, A.max_i, A.max_o, true));
    __syncthreads();
    const int r = r0 + wid;
    if (r >= n_robots) return;
    Rob<T> R = make_rob(A, r);
    WSm<T> m = wave_carve(A, reinterpret_cast<T *>(smem_raw), wid, true);
    if (A.robots[r].spring0 != A.robots[r0].spring0 || A.robots[r].weights != A.robots[r0].weights) {
        if (lane == 0) R.Wg1[0] = T(NAN);   // robots of one workgroup must share a design
        return;
    }
    wave_point_at_shared(R, sh);
    wave_bwd_robot<T, IMAX, OMAX>(A, R, m, sh, r, lane);
}

// Segments the adjoint of a robot walks from t_bwd with a step budget.
__device__ __forceinline__ int bwd_segments(int tb, int budget, int K) {
    int n = 0;
    while (tb > 0 && budget > 0) {
        const int s0 = (tb - 1) / K * K, lo = max(s0, tb - budget);
        budget -= tb - lo;
        tb = lo;
        ++n;
    }
    return n;
}

// Adjoint with the segment recompute on a second wavefront per robot
// (wavefront j: adjoint of robot j, wavefront nrob + j: its recompute; the
// wavefronts of a workgroup go to the four SIMDs in turn, so each SIMD runs
// one adjoint and one recompute wavefront), pipelined over two
// scratch buffers: while the adjoint wavefront reverses segment s from one
// buffer, the recompute wavefront folds the weight gradient of segment s + 1
// from the other and then rebuilds segment s - 1 into it.
// One workgroup barrier per segment; every wavefront runs the workgroup's
// largest segment count, idle once its robot is done (or absent), so the
// barrier counts always agree.
// The recompute keeps each step's per-vertex impulse sums with the rows, so
// the adjoint wavefront runs no spring pass of its own.
// STORED: the forward wrote every step's rows (scratch of K >= horizon rows,
// slot = step), so nothing is recomputed: the adjoint wavefront reverses the
// horizon in chunks of kRowsChunk steps and the second wavefront folds the
// weight gradient of the chunk reversed before.
constexpr int kRowsChunk = 50;

template <typename T, int IMAX, int OMAX, bool STORED>
__global__ void __launch_bounds__(512) k_wave_bwd_pipe(RArgs<T> A, int n_robots) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int iters;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int nrob = blockDim.x >> 7, rl = wid % nrob, role = wid / nrob;
    const int r0 = blockIdx.x * nrob;
    const int K = STORED ? kRowsChunk : A.K;   // steps per segment
    const int budget0 = A.max_steps > 0 ? A.max_steps : 0x7fffffff;
    const Rob<T> R0 = make_rob(A, r0);
    const size_t Rb = wave_region(A.max_n, A.max_s, A.max_i, A.max_o, true);
    const size_t Rf = wave_region(A.max_n, A.max_s, A.max_i, A.max_o, false);
    const WShared<T> sh = wave_stage(A, R0, smem_raw, true, (size_t)nrob * (Rb + Rf));
    if (threadIdx.x == 0) {
        int mx = 0;
        for (int j = 0; j < nrob && r0 + j < n_robots; ++j)
            mx = max(mx, bwd_segments(A.robots[r0 + j].t_bwd, budget0, K));
        iters = mx;
    }
    __syncthreads();
    const int r = r0 + rl;
    bool active = r < n_robots;
    if (active && (A.robots[r].spring0 != A.robots[r0].spring0 ||
                   A.robots[r].weights != A.robots[r0].weights ||
                   (STORED && A.robots[r].horizon > A.K))) {
        // robots of one workgroup must share a design (and stored rows need
        // K >= horizon rows of scratch)
        // (the forward launch over the same table reports it in A.err)
        if (role == 0 && lane == 0) make_rob(A, r).Wg1[0] = T(NAN);
        active = false;
    }
    Rob<T> R = make_rob(A, active ? r : r0);
    wave_point_at_shared(R, sh);
    T *const scr0 = R.sx, *const scr1 = R.scr2;   // scratch buffers 0 and 1
    // Outside the stored mode the reversing wavefront folds the W1 rows of the
    // last input block and the b1, W2, b2 sums of each segment as soon as it
    // has reversed it (its state rows are dead then), and the recomputing
    // wavefront folds the other input blocks one segment later.  The two write
    // disjoint entries, and every entry still sums its steps in order.
    const int isplit = ((R.I + 63) / 64 - 1) * 64;
    T *region = reinterpret_cast<T *>(smem_raw) + (size_t)rl * (Rb + Rf);
    WSm<T> m = role == 0 ? wave_carve_at(A, region, true) : wave_carve_at(A, region + Rb, false);
    int tb = active ? A.robots[r].t_bwd : 0, budget = budget0;
    const bool work = tb > 0;
    WRegs<T, 1, OMAX> w;
    LaneTopo<T> tp;
    if (work) {
        for (int q = lane; q < (A.max_i > 192 ? A.max_i : 192) + 4; q += 64) m.inb[q] = T(0);
        wave_load_weights(R, A.H, lane, w);
        lane_topo_load(R, lane, tp);
        if (role == 0) wave_bwd_begin(A, R, m, r, lane);
        else if (!STORED)
            wave_recompute<T, OMAX, true>(A, R, m, w, tp, sh.sin_tab, lane, (tb - 1) / K * K, tb);
    }
    __syncthreads();
    int buf = 0;
    int fs0 = 0, flo = 0, ftb = 0;   // segment reversed last, its weight gradient not folded yet
    for (int it = 0; it < iters; ++it) {
        const bool go = tb > 0 && budget > 0;
        const int s0 = go ? (tb - 1) / K * K : 0;
        const int lo = go ? max(s0, tb - budget) : 0;
        const int rest = go ? budget - (tb - lo) : 0;
        if constexpr (STORED) {
            // one buffer of all the steps' rows, slots = steps
            carve_scratch(R, scr0, A.K, A.H);
            if (role == 0) {
                if (go) wave_bwd_steps<T, IMAX, OMAX, true>(A, R, m, w, tp, r, lane, 0, lo, tb);
            } else if (ftb > flo) {
                wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, false), lane,
                                     flo, ftb - flo);
            }
        } else if (role == 0) {
            if (go) {
                carve_scratch(R, buf ? scr1 : scr0, K, A.H);
                wave_bwd_steps<T, IMAX, OMAX, true>(A, R, m, w, tp, r, lane, s0, lo, tb);
                wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, true), lane,
                                     lo - s0, tb - lo, isplit, 1 << 30, true);
            }
        } else {
            carve_scratch(R, buf ? scr0 : scr1, K, A.H);
            if (ftb > flo) {
                wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, false), lane,
                                     flo - fs0, ftb - flo, 0, isplit, false);
                // the fold's reads of this buffer complete before it is rebuilt
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
                wave_sync();
            }
            if (go && lo == s0 && s0 > 0 && rest > 0)
                wave_recompute<T, OMAX, true>(A, R, m, w, tp, sh.sin_tab, lane, s0 - K, s0);
        }
        fs0 = s0;
        flo = lo;
        ftb = go ? tb : lo;
        __syncthreads();
        if (go) {
            budget = rest;
            tb = lo;
            buf ^= 1;
        }
    }
    if (work && role == 0) wave_bwd_end(R, m, lane);
    if (role == 1 && ftb > flo) {
        if constexpr (STORED) {
            carve_scratch(R, scr0, A.K, A.H);
            wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, false), lane,
                                 flo, ftb - flo);
        } else {
            carve_scratch(R, buf ? scr0 : scr1, K, A.H);
            wave_seg_wgrad<OMAX>(A, R, m.J, wave_fold_cap(A.max_n, A.max_s, false), lane,
                                 flo - fs0, ftb - flo, 0, isplit, false);
        }
    }
}

// Fixed-order sum of per-robot gradient partials into per-design gradients.
// A workgroup (256 threads) takes kWrEntries entries at a time: thread (g, j)
// sums robots g, g + kWrPhases, g + 2 kWrPhases, ... of the design's list for
// entry j, and the kWrPhases partials are added in a fixed tree, so the
// result does not depend on the launch shape.
constexpr int kWrPhases = 8, kWrEntries = 32;
template <typename T>
__global__ void __launch_bounds__(256)
    k_wgrad_reduce(const int32_t *lptr, const int32_t *list, const int64_t *roff,
                   const int64_t *doff, const int64_t *dsize, const T *wg, T *out) {
    static_assert(kWrPhases * kWrEntries == 256, "one workgroup of 256 threads");
    __shared__ T part[kWrPhases][kWrEntries];
    const int d = blockIdx.y;
    const int64_t n = dsize[d];
    const int j = threadIdx.x % kWrEntries, g = threadIdx.x / kWrEntries;
    const int q0 = lptr[d], q1 = lptr[d + 1];
    for (int64_t e0 = (int64_t)blockIdx.x * kWrEntries; e0 < n;
         e0 += (int64_t)gridDim.x * kWrEntries) {
        const int64_t e = e0 + j;
        T s = T(0);
        if (e < n) {
#pragma unroll 4
            for (int q = q0 + g; q < q1; q += kWrPhases) s += wg[roff[list[q]] + e];
        }
        part[g][j] = s;
        __syncthreads();
        if (g == 0 && e < n) {
            const T a = (part[0][j] + part[1][j]) + (part[2][j] + part[3][j]);
            const T b = (part[4][j] + part[5][j]) + (part[6][j] + part[7][j]);
            out[doff[d] + e] = a + b;
        }
        __syncthreads();
    }
}

}  // namespace dl

using namespace dl;

template <typename T>
static RArgs<T> make_rargs(const dl_rollout *r) {
    RArgs<T> A;
    A.robots = r->robots;
    A.sa = r->spring_a; A.sb = r->spring_b; A.sact = r->spring_act;
    A.L0 = (const T *)r->spring_L0; A.k = (const T *)r->spring_k; A.coef = (const T *)r->spring_coeff;
    A.vptr = r->vert_ptr; A.vidx = r->vert_idx; A.aptr = r->act_ptr; A.aidx = r->act_idx;
    A.W = (const T *)r->weights; A.Wg = (T *)r->wgrad;
    A.state = (T *)r->state; A.adj = (T *)r->adjoint; A.ckpt = (T *)r->ckpt;
    A.scratch = (T *)r->scratch;
    A.targets = (const T *)r->targets;
    A.loss = (T *)r->loss; A.loss_grad = (const T *)r->loss_grad;
    A.traj = (T *)r->traj; A.traj_grad = (const T *)r->traj_grad;
    A.err = r->err;
    A.P = make_phys<T>(r->phys);
    A.L = make_layout(r->lay);
    A.H = r->n_hidden; A.K = r->ckpt_every; A.max_steps = r->max_steps; A.act = r->act;
    A.loss_kind = r->loss_kind; A.window = r->loss_window; A.store_traj = r->store_traj;
    A.max_n = r->max_n; A.max_s = r->max_s; A.max_i = r->max_i; A.max_o = r->max_o;
    A.springs_in_lds = 1;
    A.topo_in_lds = 0;
    A.csr_in_lds = 0;
    A.w_in_lds = 0;
    return A;
}

static int check_rollout(const dl_rollout *r, bool bwd) {
    DL_CHECK_ARG(r != nullptr, "rollout descriptor is NULL");
    DL_CHECK_ARG(r->dtype == DL_F32 || r->dtype == DL_F64, "rollout: bad dtype %d", r->dtype);
    DL_CHECK_ARG(r->act == DL_ACT_SIN || r->act == DL_ACT_TANH, "rollout: bad activation");
    DL_CHECK_ARG(r->n_robots > 0 && r->n_robots <= 65535 * 1024, "rollout: bad robot count");
    DL_CHECK_ARG(r->n_hidden > 0 && r->n_hidden <= HMAX, "rollout: hidden width must be 1..%d", HMAX);
    DL_CHECK_ARG(r->ckpt_every > 0, "rollout: checkpoint interval must be positive");
    DL_CHECK_ARG(r->block >= 64 && r->block % 64 == 0 &&
                     (r->block <= 512 || (r->block == 1024 && r->dtype == DL_F32)),
                 "rollout: block must be a multiple of 64 in [64, 512] (or 1024 in FP32)");
    DL_CHECK_ARG(r->max_n > 0 && r->max_s >= 0 && r->max_i > 0 && r->max_o > 0,
                 "rollout: bad size maxima");
    DL_CHECK_ARG(r->lds_policy >= 0 && r->lds_policy <= 4, "rollout: bad LDS policy");
    DL_CHECK_ARG(r->variant >= 0 && r->variant <= 5, "rollout: bad kernel variant");
    DL_CHECK_ARG(r->lay.n_tv >= 0 && r->lay.n_tv <= 3 && r->lay.run_period > 0, "rollout: bad layout");
    DL_CHECK_ARG(r->loss_kind == 0 || (r->loss_kind == 1 && r->loss_window > 0),
                 "rollout: bad loss configuration");
    DL_CHECK_ARG(r->robots && r->spring_a && r->spring_b && r->spring_act && r->spring_L0 &&
                     r->spring_k && r->spring_coeff && r->vert_ptr && r->vert_idx && r->act_ptr &&
                     r->act_idx && r->weights && r->state && r->ckpt && r->targets,
                 "rollout: NULL buffer");
    DL_CHECK_ARG(r->loss_kind == 0 || r->loss, "rollout: loss buffer required");
    if (bwd)
        DL_CHECK_ARG(r->adjoint && r->scratch && r->wgrad && (r->loss_kind == 0 || r->loss_grad),
                     "rollout backward: adjoint, scratch, wgrad and loss_grad buffers required");
    return 0;
}

static const size_t kLdsLimit = 160 * 1024;

template <int IM>
static int launch_wave_im(const dl_rollout *r, const RArgs<float> &A, bool bwd, bool pipe, bool stored,
                          dim3 grid, dim3 block, size_t sh, hipStream_t st) {
    const void *kf = pipe ? (stored ? (const void *)k_wave_bwd_pipe<float, IM, 8, true>
                                    : (const void *)k_wave_bwd_pipe<float, IM, 8, false>)
                     : bwd ? (const void *)k_wave_bwd<float, IM, 8>
                           : (stored ? (const void *)k_wave_fwd<float, 8, true>
                                     : (const void *)k_wave_fwd<float, 8, false>);
    DL_CHECK_HIP(ensure_dynamic_lds(kf, sh));
    if (pipe && stored)
        hipLaunchKernelGGL((k_wave_bwd_pipe<float, IM, 8, true>), grid, block, sh, st, A, r->n_robots);
    else if (pipe)
        hipLaunchKernelGGL((k_wave_bwd_pipe<float, IM, 8, false>), grid, block, sh, st, A, r->n_robots);
    else if (bwd)
        hipLaunchKernelGGL((k_wave_bwd<float, IM, 8>), grid, block, sh, st, A, r->n_robots);
    else if (stored)
        hipLaunchKernelGGL((k_wave_fwd<float, 8, true>), grid, block, sh, st, A, r->n_robots);
    else
        hipLaunchKernelGGL((k_wave_fwd<float, 8, false>), grid, block, sh, st, A, r->n_robots);
    DL_CHECK_LAUNCH();
    return 0;
}

// stored: the forward writes every step's adjoint rows and the pipelined
// adjoint reads them (no segment recompute).
template <typename T>
static int launch_wave(const dl_rollout *r, RArgs<T> A, bool bwd, bool pipelined, bool stored,
                       hipStream_t st) {
    if constexpr (sizeof(T) != 4) {
        DL_CHECK_ARG(false, "rollout: wave kernel is FP32 only");
    } else {
        const int waves = r->block / 64;   // robots per workgroup (all of one design)
        const int n_tab = r->lay.run_period * r->lay.n_sin;
        const bool pipe = bwd && pipelined;
        const size_t sh = pipe ? wave_pipe_smem_bytes(r->max_n, r->max_s, r->max_i, r->max_o, n_tab,
                                                      waves)
                               : wave_smem_bytes(r->max_n, r->max_s, r->max_i, r->max_o, bwd, n_tab,
                                                 waves);
        const dim3 grid((r->n_robots + waves - 1) / waves), block((pipe ? 128 : 64) * waves);
        if (r->max_i <= 64) return launch_wave_im<64>(r, A, bwd, pipe, stored, grid, block, sh, st);
        if (r->max_i <= 128) return launch_wave_im<128>(r, A, bwd, pipe, stored, grid, block, sh, st);
        return launch_wave_im<192>(r, A, bwd, pipe, stored, grid, block, sh, st);
    }
    return 0;
}

template <typename T>
static int launch_rollout(const dl_rollout *r, bool bwd, hipStream_t st) {
    RArgs<T> A = make_rargs<T>(r);
    // LDS placement: policy 0 = auto (topology when it fits in 64 KiB, W1 too
    // when the total stays within 80 KiB), 1 = topology only, 2 = topology +
    // W1, 3 = nothing staged, 4 = as much as fits in the whole LDS (spring
    // buffers, then topology, then W1; one workgroup per CU for big robots).
    // Under policies 0 and 4 a robot whose topology does not fit keeps its
    // vertex / actuator spring lists in LDS when they fit.
    const int pol = r->lds_policy;
    auto bytes = [&](bool springs, bool topo, bool wl, bool csr = false) {
        return smem_bytes<T>(r->max_n, r->max_s, r->max_i, r->max_o, bwd, springs, topo, wl, csr);
    };
    const size_t cap = pol == 4 ? kLdsLimit : 64 * 1024;
    bool springs = bytes(true, false, false) <= cap;
    bool topo = pol != 3, wl = pol == 2 || pol == 0, csr = false;
    if (pol == 0 || pol == 4) {
        topo = bytes(springs, true, false) <= cap;
        wl = topo && r->n_hidden == HMAX &&
             bytes(springs, true, true) <= (pol == 4 ? kLdsLimit : 80 * 1024);
        csr = bwd && !topo && bytes(springs, false, false, true) <= cap;
    }
    // wave-per-robot kernel: FP32, hidden width 64, inputs <= 192, outputs <= 8
    const int variant = r->variant;
    const int n_tab = r->lay.run_period * r->lay.n_sin;
    const bool wave_ok = sizeof(T) == 4 && r->n_hidden == HMAX && r->max_i <= 192 &&
                         r->max_o <= 8 && r->max_n <= 64 && 7 * r->max_s >= HMAX &&
                         wave_fold_cap(r->max_n, r->max_s, true) >= kFoldStepReals && r->lay.n_tv >= 1 &&
                         (r->block == 64 || r->block == 256) &&
                         wave_smem_bytes(r->max_n, r->max_s, r->max_i, r->max_o, bwd, n_tab,
                                         r->block / 64) <= kLdsLimit;
    DL_CHECK_ARG((variant != 2 && variant != 3) || wave_ok,
                 "rollout: the wave-per-robot kernel needs FP32, H=64, I<=192, O<=8, block 64 or 256");
    // pipelined adjoint: a second wavefront per robot (the scratch holds two
    // segment buffers, see dl_scratch_size)
    const bool pipe_ok = wave_ok && r->scratch != nullptr && 3 * r->max_s >= HMAX &&
                         wave_fold_cap(r->max_n, r->max_s, false) >= kFoldStepReals &&
                         wave_pipe_smem_bytes(r->max_n, r->max_s, r->max_i, r->max_o, n_tab,
                                              r->block / 64) <= kLdsLimit;
    DL_CHECK_ARG((variant != 3 && variant != 4) || pipe_ok,
                 "rollout: the pipelined wave adjoint does not fit in LDS or has no scratch");
    // variant 4 keeps every step's rows: the forward needs the scratch too
    DL_CHECK_ARG(variant != 4 || r->scratch != nullptr, "rollout: stored rows need the scratch buffer");
    if (variant == 4) return launch_wave<T>(r, A, bwd, true, true, st);
    DL_CHECK_ARG(variant != 5 || r->scratch != nullptr, "rollout: stored rows need the scratch buffer");
    if (variant == 3 || (variant == 0 && pipe_ok)) return launch_wave<T>(r, A, bwd, true, false, st);
    if (variant == 2 || (variant == 0 && wave_ok)) return launch_wave<T>(r, A, bwd, false, false, st);
    size_t sh = bytes(springs, topo, wl, csr);
    DL_CHECK_ARG(sh <= kLdsLimit, "rollout: robot too large for LDS (%zu bytes)", sh);
    DL_CHECK_ARG(springs || r->scratch != nullptr,
                 "rollout: robot spills springs to scratch; scratch required");
    A.springs_in_lds = springs;
    A.topo_in_lds = topo;
    A.csr_in_lds = topo || csr;
    A.w_in_lds = wl;
    const bool big = r->block > 256, rows = variant == 5;
    const dim3 grid(r->n_robots), block(r->block);
    if constexpr (std::is_same_v<T, float>) {
        // 16 wavefronts per robot (FP32 only: the register budget per lane is
        // 128 at this size)
        if (r->block > 512) {
            const void *kh = bwd ? (rows ? (const void *)k_rollout_bwd<T, 1024, true>
                                         : (const void *)k_rollout_bwd<T, 1024, false>)
                                 : (rows ? (const void *)k_rollout_fwd<T, 1024, true>
                                         : (const void *)k_rollout_fwd<T, 1024, false>);
            DL_CHECK_HIP(ensure_dynamic_lds(kh, sh));
            if (bwd && rows) hipLaunchKernelGGL((k_rollout_bwd<T, 1024, true>), grid, block, sh, st, A);
            else if (bwd) hipLaunchKernelGGL((k_rollout_bwd<T, 1024, false>), grid, block, sh, st, A);
            else if (rows) hipLaunchKernelGGL((k_rollout_fwd<T, 1024, true>), grid, block, sh, st, A);
            else hipLaunchKernelGGL((k_rollout_fwd<T, 1024, false>), grid, block, sh, st, A);
            DL_CHECK_LAUNCH();
            return 0;
        }
    }
    const void *kf =
        bwd ? (big ? (rows ? (const void *)k_rollout_bwd<T, 512, true> : (const void *)k_rollout_bwd<T, 512, false>)
                   : (rows ? (const void *)k_rollout_bwd<T, 256, true> : (const void *)k_rollout_bwd<T, 256, false>))
            : (big ? (rows ? (const void *)k_rollout_fwd<T, 512, true> : (const void *)k_rollout_fwd<T, 512, false>)
                   : (rows ? (const void *)k_rollout_fwd<T, 256, true> : (const void *)k_rollout_fwd<T, 256, false>));
    DL_CHECK_HIP(ensure_dynamic_lds(kf, sh));
    if (bwd && big && rows) hipLaunchKernelGGL((k_rollout_bwd<T, 512, true>), grid, block, sh, st, A);
    else if (bwd && big) hipLaunchKernelGGL((k_rollout_bwd<T, 512, false>), grid, block, sh, st, A);
    else if (bwd && rows) hipLaunchKernelGGL((k_rollout_bwd<T, 256, true>), grid, block, sh, st, A);
    else if (bwd) hipLaunchKernelGGL((k_rollout_bwd<T, 256, false>), grid, block, sh, st, A);
    else if (big && rows) hipLaunchKernelGGL((k_rollout_fwd<T, 512, true>), grid, block, sh, st, A);
    else if (big) hipLaunchKernelGGL((k_rollout_fwd<T, 512, false>), grid, block, sh, st, A);
    else if (rows) hipLaunchKernelGGL((k_rollout_fwd<T, 256, true>), grid, block, sh, st, A);
    else hipLaunchKernelGGL((k_rollout_fwd<T, 256, false>), grid, block, sh, st, A);
    DL_CHECK_LAUNCH();
    return 0;
}

extern "C" {

int64_t dl_weights_size(int32_t I, int32_t H, int32_t O) {
    return (2 * (int64_t)H * I + H + 2 * (int64_t)O * H + O + 15) / 16 * 16;
}

int64_t dl_wgrad_size(int32_t I, int32_t H, int32_t O) {
    return (int64_t)H * I + H + (int64_t)O * H + O;
}

int64_t dl_scratch_size(int32_t N, int32_t S, int32_t I, int32_t H, int32_t O, int32_t K) {
    return 2 * (int64_t)K * (9 * (int64_t)N + I + 3 * (int64_t)H + 3 * (int64_t)O + 6) +
           10 * (int64_t)S + 4;
}

int32_t dl_ms_rollout_forward(const dl_rollout *r, void *stream) {
    if (int e = check_rollout(r, false)) return e;
    if (r->dtype == DL_F32) return launch_rollout<float>(r, false, (hipStream_t)stream);
    return launch_rollout<double>(r, false, (hipStream_t)stream);
}

int32_t dl_ms_rollout_backward(const dl_rollout *r, void *stream) {
    if (int e = check_rollout(r, true)) return e;
    if (r->dtype == DL_F32) return launch_rollout<float>(r, true, (hipStream_t)stream);
    return launch_rollout<double>(r, true, (hipStream_t)stream);
}

int32_t dl_wgrad_reduce(int32_t dtype, int32_t n_robots, const int32_t *lptr, const int32_t *list,
                        int32_t n_designs, const int64_t *roff, const int64_t *doff,
                        const int64_t *dsize, const void *wg, void *out, void *stream) {
    DL_CHECK_ARG(n_robots > 0 && n_designs > 0 && n_designs <= 65535 && lptr && list && roff &&
                     doff && dsize && wg && out,
                 "dl_wgrad_reduce: bad arguments");
    // one design: enough workgroups for its entries in one pass; many
    // designs: fewer per design
    const int gx = n_designs == 1 ? 512 : std::max(16, 8192 / n_designs);
    if (dtype == DL_F32)
        hipLaunchKernelGGL(k_wgrad_reduce<float>, dim3(gx, n_designs), dim3(256), 0,
                           (hipStream_t)stream, lptr, list, roff, doff, dsize, (const float *)wg,
                           (float *)out);
    else if (dtype == DL_F64)
        hipLaunchKernelGGL(k_wgrad_reduce<double>, dim3(gx, n_designs), dim3(256), 0,
                           (hipStream_t)stream, lptr, list, roff, doff, dsize, (const double *)wg,
                           (double *)out);
    else
        DL_CHECK_ARG(false, "dl_wgrad_reduce: bad dtype");
    DL_CHECK_LAUNCH();
    return 0;
}

}  // extern "C"
