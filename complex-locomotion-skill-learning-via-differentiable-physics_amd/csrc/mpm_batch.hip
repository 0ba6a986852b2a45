// Batched MLS-MPM training rollout: B robots, each on its own grid, advanced
// and differentiated a segment of steps per host call (dl_mpm_forward /
// dl_mpm_backward), so no per-step work crosses the C ABI.
//
// Follows multitask/solver_mpm.py: p2g (:83-124), grid_op (:126-144), g2p
// (:146-165), compute_center (:167-173) and advance_core(_grad) (:175-190),
// with the quadruped controller (COM and per-group sensors, MLP) in front.
//
// Layouts (per robot k):
//   particle slots   x, v [S][B][D][P];  C, F [S][B][D*D][P]
//   grids            [B][n^D][4] records: grid = (momentum (D), mass),
//                    grid_vo = velocity, grid_vo_grad = its adjoint,
//                    grid_grad = adjoint of (momentum, mass)
//   active blocks    flag/list [B][nblk], count [B]; a block is 4^3 nodes
//                    (3D) or 8^2 (2D) -- one wavefront per block.
//
// Grid sparsity: P2G appends every block its particles touch to the robot's
// active list; the grid passes visit only listed blocks and clear what they
// consume, so grid memory is zero between steps and no dense clear or dense
// pass runs.  Invariants between calls: grid, grid_vo_grad, flags, count and
// act_grad are zero.
//
// P2G and the G2P adjoint accumulate a workgroup's particles into an LDS
// tile covering their stencils' bounding box, then commit each touched node
// with one global atomic per component (direct global atomics when the box
// exceeds the tile).  Atomic summation order is the run-to-run variation.

#include "mpm_core.hpp"

#include <algorithm>
#include <climits>

namespace dl {

constexpr int TILE_MAX = DL_MPM_TILE_MAX;       // largest LDS tile, nodes
constexpr int TILE_CAP = DL_MPM_TILE_CAP;       // tiles listed per block
constexpr int PB = DL_MPM_PB;                   // particles per workgroup
constexpr int NWAVE = PB / 64;
constexpr int MAX_TILE_BLOCKS = 256;            // blocks a stored tile may overlap

// Stored scatter tiles of one robot: tile w at tiles + w*tn*4 with
// its node box (lo[3], ext[3]) at hdr + 8w; per block the ids of the tiles
// overlapping it (count, ids[TILE_CAP]).
template <typename T> struct TileBuf {
    T *tiles;
    int32_t *hdr, *cnt, *ids;
    int wpr;                                    // workgroups (tiles) per robot
    int tn;                                     // tile capacity, nodes
    int store;                                  // commit tiles by store + gather
    __device__ TileBuf robot(int k, int nblk) const {
        return {tiles + (size_t)k * wpr * tn * 4, hdr + (size_t)k * wpr * 8,
                cnt + (size_t)k * nblk, ids + (size_t)k * nblk * TILE_CAP, wpr, tn, store};
    }
};

template <int D> struct BlkE { static constexpr int E = D == 3 ? 4 : 8; };

struct Blk {
    int32_t *flag, *list, *count;
    int nb, nblk;
};

template <int D>
__host__ __device__ __forceinline__ size_t gsize(int n) {
    size_t g = 1;
    for (int d = 0; d < D; ++d) g *= (size_t)n;
    return g;
}

// ---------------------------------------------------------------- AoS grid io
template <typename T> __device__ __forceinline__ void ld4(const T *p, T o[4]);
template <> __device__ __forceinline__ void ld4<float>(const float *p, float o[4]) {
    const float4 v = *reinterpret_cast<const float4 *>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
template <> __device__ __forceinline__ void ld4<double>(const double *p, double o[4]) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    const double2 b = *reinterpret_cast<const double2 *>(p + 2);
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
template <typename T> __device__ __forceinline__ void st4(T *p, const T o[4]);
template <> __device__ __forceinline__ void st4<float>(float *p, const float o[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
}
template <> __device__ __forceinline__ void st4<double>(double *p, const double o[4]) {
    *reinterpret_cast<double2 *>(p) = make_double2(o[0], o[1]);
    *reinterpret_cast<double2 *>(p + 2) = make_double2(o[2], o[3]);
}

template <typename T>
__device__ __forceinline__ T wsel(int i, T a, T b, T c) { return i == 0 ? a : (i == 1 ? b : c); }

template <typename T>
__device__ __forceinline__ void gadd(T *p, T v) { unsafeAtomicAdd(p, v); }

// Bounding box of the base cells of the workgroup's in-grid particles.
// Returns false (uniformly) when the workgroup has none.
template <int D>
__device__ __forceinline__ bool block_bbox(bool valid, const int base[D], int lo[D], int ext[D],
                                           int *red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        int mn = valid ? base[d] : INT_MAX, mx = valid ? base[d] : INT_MIN;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = min(mn, __shfl_xor(mn, off, 64));
            mx = max(mx, __shfl_xor(mx, off, 64));
        }
        if (lane == 0) {
            red[(2 * d) * NWAVE + wid] = mn;
            red[(2 * d + 1) * NWAVE + wid] = mx;
        }
    }
    __syncthreads();
    bool any = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        int mn = INT_MAX, mx = INT_MIN;
#pragma unroll
        for (int w = 0; w < NWAVE; ++w) {
            mn = min(mn, red[(2 * d) * NWAVE + w]);
            mx = max(mx, red[(2 * d + 1) * NWAVE + w]);
        }
        any = any && mn <= mx;
        lo[d] = mn;
        ext[d] = mx - mn + 3;
    }
    __syncthreads();
    return any;
}

// Append every block overlapped by the node box [lo, lo + ext) to robot k's
// active list (first marker wins; marking a block with no mass is harmless).
template <int D>
__device__ __forceinline__ void mark_box(const Blk &bk, int k, const int lo[D], const int ext[D]) {
    constexpr int E = BlkE<D>::E;
    int b0[D], nbt[D], tot = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        b0[d] = lo[d] / E;
        nbt[d] = (lo[d] + ext[d] - 1) / E - b0[d] + 1;
        tot *= nbt[d];
    }
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        int rem = i, c[D];
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            c[d] = b0[d] + rem % nbt[d];
            rem /= nbt[d];
        }
        int bid = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) bid = bid * bk.nb + c[d];
        int32_t *f = bk.flag + (size_t)k * bk.nblk + bid;
        if (*reinterpret_cast<volatile int32_t *>(f) == 0 && atomicCAS(f, 0, 1) == 0) {
            const int slot = atomicAdd(bk.count + k, 1);
            bk.list[(size_t)k * bk.nblk + slot] = bid;
        }
    }
}

// Node of lane `lane` in block `bid`; sets bnd for sticky boundary nodes.
template <int D, typename T>
__device__ __forceinline__ int block_node(int bid, int lane, int nb, const MP<T> &p, bool &bnd,
                                          int c[D]) {
    constexpr int E = BlkE<D>::E;
    int rem = bid, lr = lane;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
        c[d] = (rem % nb) * E + lr % E;
        rem /= nb;
        lr /= E;
    }
    int node = 0;
    bnd = false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        node = node * p.n + c[d];
        bnd = bnd || c[d] < p.bound || c[d] > p.n - p.bound;
    }
    return node;
}

// Wave-aggregated accumulation of per-particle values into LDS slots keyed
// by a small integer (sensor group, actuator id): one atomic per distinct
// key in the wave.  All lanes of the wave must call it.
template <typename T, int NV>
__device__ __forceinline__ void keyed_wave_add(bool valid, int key, const T (&v)[NV], T *dst,
                                               int stride) {
    const int lane = threadIdx.x & 63;
    unsigned long long pend = __ballot(valid);
    while (pend) {
        const int leader = __ffsll((long long)pend) - 1;
        const int kl = __shfl(key, leader, 64);
        const bool mine = valid && key == kl;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const T s = wave_sum(mine ? v[i] : T(0));
            if (lane == leader) atomicAdd(dst + kl * stride + i, s);
        }
        pend &= ~__ballot(mine);
    }
}

// Store a workgroup's whole LDS tile (nodes <= tile capacity) with plain
// coalesced writes and list it under every block its box overlaps; the
// grid pass then gathers the listed tiles, so no global atomics are needed.
// A block whose list is full gets this tile's nodes by global atomics
// instead (into gk).  Returns false, storing nothing, when the box overlaps
// more than MAX_TILE_BLOCKS blocks (the caller then commits by atomics).
// Workgroup-uniform; called after the tile is complete (post barrier).
template <int D, typename T>
__device__ __forceinline__ bool store_tile(const T *tile, int nodes, const int lo[D],
                                           const int ext[D], const TileBuf<T> &tb, int nb,
                                           T *gk, int n) {
    constexpr int E = BlkE<D>::E;
    int b0[D], nbt[D], tot = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        b0[d] = lo[d] / E;
        nbt[d] = (lo[d] + ext[d] - 1) / E - b0[d] + 1;
        tot *= nbt[d];
    }
    if (tot > MAX_TILE_BLOCKS) return false;
    __shared__ unsigned int full_mask[MAX_TILE_BLOCKS / 32];
    __shared__ int any_full;
    const int w = blockIdx.x;
    T *dst = tb.tiles + (size_t)w * tb.tn * 4;
    for (int i = threadIdx.x; i < nodes; i += blockDim.x) st4(dst + (size_t)i * 4, tile + i * 4);
    if (threadIdx.x < D) {
        tb.hdr[w * 8 + threadIdx.x] = lo[threadIdx.x];
        tb.hdr[w * 8 + 3 + threadIdx.x] = ext[threadIdx.x];
    }
    if (threadIdx.x < MAX_TILE_BLOCKS / 32) full_mask[threadIdx.x] = 0u;
    if (threadIdx.x == 0) any_full = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        int rem = i, bid = 0, c[D];
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {
            c[d] = b0[d] + rem % nbt[d];
            rem /= nbt[d];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) bid = bid * nb + c[d];
        const int slot = atomicAdd(tb.cnt + bid, 1);
        if (slot < TILE_CAP) {
            tb.ids[(size_t)bid * TILE_CAP + slot] = w;
        } else {
            atomicOr(full_mask + (i >> 5), 1u << (i & 31));
            any_full = 1;
        }
    }
    __syncthreads();
    if (any_full) {
        for (int i = threadIdx.x; i < nodes; i += blockDim.x) {
            int rem = i, node = 0, c[D], bl = 0;
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                c[d] = lo[d] + rem % ext[d];
                rem /= ext[d];
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                node = node * n + c[d];
                bl = bl * nbt[d] + c[d] / E - b0[d];
            }
            if (full_mask[bl >> 5] & (1u << (bl & 31))) {
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    const T r = tile[i * 4 + cc];
                    if (r != T(0)) gadd(gk + (size_t)node * 4 + cc, r);
                }
            }
        }
    }
    return true;
}

// Sum of the stored tiles listed under block `bid` at node coordinates c.
template <int D, typename T>
__device__ __forceinline__ void gather_tiles(const TileBuf<T> &tb, int bid, const int c[D],
                                             T r[4]) {
    const int nt = min(tb.cnt[bid], TILE_CAP);
    for (int s = 0; s < nt; ++s) {
        const int w = tb.ids[(size_t)bid * TILE_CAP + s];
        const int32_t *h = tb.hdr + w * 8;
        int li = 0;
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int q = c[d] - h[d];
            in = in && q >= 0 && q < h[3 + d];
            li = li * h[3 + d] + q;
        }
        if (in) {
            T t4[4];
            ld4(tb.tiles + (size_t)w * tb.tn * 4 + (size_t)li * 4, t4);
#pragma unroll
            for (int cc = 0; cc < 4; ++cc) r[cc] += t4[cc];
        }
    }
}

// ------------------------------------------------------------ DPP row shifts
// Lane i reads lane i-N (shr) or i+N (shl) of its 16-lane row; lanes without
// a source read 0.  VALU only (no LDS traffic).
template <int N> __device__ __forceinline__ int dpp_shr_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x110 + N, 0xF, 0xF, true);
}
template <int N> __device__ __forceinline__ int dpp_shl_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 0x100 + N, 0xF, 0xF, true);
}
template <int N> __device__ __forceinline__ float dpp_shr(float v) {
    return __int_as_float(dpp_shr_i<N>(__float_as_int(v)));
}
template <int N> __device__ __forceinline__ double dpp_shr(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shr_i<N>((int)(b & 0xffffffffll)), hi = dpp_shr_i<N>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Runs of consecutive lanes (within a 16-lane row) holding the same key --
// particles of one grid cell, since particles are sorted by cell.  dist =
// lanes before this one in its run; tail = last lane of its run.
struct Seg {
    int dist;
    bool tail;
};
// The lane shifts must run with every lane active (a disabled source lane
// reads 0), so they are evaluated unconditionally, never under a
// short-circuit condition.
__device__ __forceinline__ Seg lane_runs(int key) {
    const int r = threadIdx.x & 15;
    const int prev = dpp_shr_i<1>(key), next = dpp_shl_i<1>(key);
    const bool head = (r == 0) | (prev != key);
    Seg sg;
    sg.tail = (r == 15) | (next != key);
    int s = head ? r : 0;
    s = max(s, dpp_shr_i<1>(s));
    s = max(s, dpp_shr_i<2>(s));
    s = max(s, dpp_shr_i<4>(s));
    s = max(s, dpp_shr_i<8>(s));
    sg.dist = r - s;
    return sg;
}

// Inclusive segmented sum over a run; the tail lane ends with the run total.
// Step n adds the value n lanes down when the run reaches that far
// (dist >= n); the 0/1 factors m[n] are per lane.  In FP32 each step is one
// multiply-add whose first operand is the shifted value, which the compiler
// folds into a single DPP instruction.
template <typename T> struct RunMask {
    T m1, m2, m4, m8;
};
template <typename T>
__device__ __forceinline__ RunMask<T> run_mask(int dist) {
    return {dist >= 1 ? T(1) : T(0), dist >= 2 ? T(1) : T(0), dist >= 4 ? T(1) : T(0),
            dist >= 8 ? T(1) : T(0)};
}
template <typename T>
__device__ __forceinline__ T run_sum(T v, const RunMask<T> &m) {
    v = fma(dpp_shr<1>(v), m.m1, v);
    v = fma(dpp_shr<2>(v), m.m2, v);
    v = fma(dpp_shr<4>(v), m.m4, v);
    v = fma(dpp_shr<8>(v), m.m8, v);
    return v;
}

// Run sums of NC channels.  The compiler emits each step above as a DPP move
// plus an FMA (the DPP combiner does not fold into v_fmac, whose accumulator
// is tied); in FP32 the 3- and 4-channel forms issue each step as one
// v_fmac_f32 whose first operand is the row-shifted register itself
// (v[i] += m_n[i] * v[i - n], rows without a source lane read 0, as above):
// the same single-rounding FMA, so the results are bit for bit those of
// run_sum.  The channels interleave, so a register is read through DPP at
// least two instructions after its last write (the two wait states a DPP
// read needs after a VALU write); the leading s_nop 4 covers the five that
// follow a write of EXEC.  Every lane must be active.
template <typename T, int NC>
__device__ __forceinline__ void run_sums(T (&v)[NC], const RunMask<T> &m) {
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = run_sum(v[c], m);
}
#define DL_FMAC_ROW(R, M, N) \
    "v_fmac_f32_dpp %" #R ", %" #R ", %" #M " row_shr:" #N " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
template <>
__device__ __forceinline__ void run_sums<float, 4>(float (&v)[4], const RunMask<float> &m) {
    asm volatile("s_nop 4\n"
                 DL_FMAC_ROW(0, 4, 1) DL_FMAC_ROW(1, 4, 1) DL_FMAC_ROW(2, 4, 1) DL_FMAC_ROW(3, 4, 1)
                 DL_FMAC_ROW(0, 5, 2) DL_FMAC_ROW(1, 5, 2) DL_FMAC_ROW(2, 5, 2) DL_FMAC_ROW(3, 5, 2)
                 DL_FMAC_ROW(0, 6, 4) DL_FMAC_ROW(1, 6, 4) DL_FMAC_ROW(2, 6, 4) DL_FMAC_ROW(3, 6, 4)
                 DL_FMAC_ROW(0, 7, 8) DL_FMAC_ROW(1, 7, 8) DL_FMAC_ROW(2, 7, 8) DL_FMAC_ROW(3, 7, 8)
                 : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3])
                 : "v"(m.m1), "v"(m.m2), "v"(m.m4), "v"(m.m8));
}
template <>
__device__ __forceinline__ void run_sums<float, 3>(float (&v)[3], const RunMask<float> &m) {
    asm volatile("s_nop 4\n"
                 DL_FMAC_ROW(0, 3, 1) DL_FMAC_ROW(1, 3, 1) DL_FMAC_ROW(2, 3, 1)
                 DL_FMAC_ROW(0, 4, 2) DL_FMAC_ROW(1, 4, 2) DL_FMAC_ROW(2, 4, 2)
                 DL_FMAC_ROW(0, 5, 4) DL_FMAC_ROW(1, 5, 4) DL_FMAC_ROW(2, 5, 4)
                 DL_FMAC_ROW(0, 6, 8) DL_FMAC_ROW(1, 6, 8) DL_FMAC_ROW(2, 6, 8)
                 : "+v"(v[0]), "+v"(v[1]), "+v"(v[2])
                 : "v"(m.m1), "v"(m.m2), "v"(m.m4), "v"(m.m8));
}
#undef DL_FMAC_ROW

// Values scattered to stencil node base + off: channel c is
// W(off) * (b[c] + sum_d L[c][d] * off[d]) for c < NL and W(off) * b[c] for
// the rest, W(off) the product of the particle's quadratic B-spline weights.
template <int D, typename T, int NC> struct LinVal {
    T b[NC];
    T L[NC][D];
};

// Run totals of every stencil offset.  The offsets run as an outer loop over
// the axes before the last (unrolled U times) and an unrolled inner loop over
// the last axis, which advances the linear values by L[c][D-1].  For offset
// o = 3 * oo + oz, emit(o, off, v) is called on every lane, v holding the
// run total on the run's tail lane; mine(off) masks a lane's contribution.
// All lanes must be active (lane shifts).
template <int D, typename T, int NC, int NL, int U, typename Mine, typename Emit>
__device__ __forceinline__ void run_offsets(const Stencil<D, T> &s, const T wz[3],
                                            const LinVal<D, T, NC> &z, const RunMask<T> &rm,
                                            Mine &&mine, Emit &&emit) {
    constexpr int KO = D == 2 ? 3 : 9;
#pragma unroll U
    for (int oo = 0; oo < KO; ++oo) {
        int off[D];
        unflat<D - 1>(oo, off);
        T Wo = T(1), m[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) m[c] = z.b[c];
#pragma unroll
        for (int d = 0; d < D - 1; ++d) {
            Wo *= wsel(off[d], s.w[0][d], s.w[1][d], s.w[2][d]);
#pragma unroll
            for (int c = 0; c < NL; ++c) m[c] += z.L[c][d] * T(off[d]);
        }
        Wo = mine(off) ? Wo : T(0);
#pragma unroll
        for (int oz = 0; oz < 3; ++oz) {
            off[D - 1] = oz;
            const T W = Wo * wz[oz];
            T v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) v[c] = W * m[c];
            run_sums<T, NC>(v, rm);
            emit(3 * oo + oz, off, v);
#pragma unroll
            for (int c = 0; c < NL; ++c) m[c] += z.L[c][D - 1];
        }
    }
}

// Scatter-add NC values per stencil node of the workgroup's particles into
// robot grid gk ([n^D][4] records).  For each stencil offset the particles
// of one cell (a run of lanes) are summed in registers first (run_offsets),
// so memory sees one contribution per run rather than per particle.  Two
// ways to combine the runs' contributions, chosen per workgroup:
//
//  * staged (no LDS atomics per contribution): each run's tail lane stores
//    its K totals to stage[run][o] and links the run into a per-cell list
//    (one LDS exchange per run); then each thread takes nodes of the box
//    and sums, over its 3^D neighbour cells, the staged totals of their
//    runs, committing one global atomic per nonzero component.  Needs the
//    stage (runs x K x 4 reals), the lists and the padded cell heads to fit
//    the LDS buffer (tile, tb.tn x 4 reals), and atomic commits.
//  * dense tile: the node box (lo, ext) is cut into x-slabs of at most tb.tn
//    nodes, each accumulated in LDS with atomics and committed with one
//    global atomic per nonzero component (or stored, tb.store); a box whose
//    single x-plane exceeds the tile is scattered with direct global atomics.
//
// Every thread of the workgroup must call it (uniform control flow, all
// lanes active for the lane shifts).
template <int D, typename T, int NC, int NL, int U>
__device__ __forceinline__ void tile_scatter(bool valid, const Stencil<D, T> &s, const int lo[D],
                                             const int ext[D], T *tile, T *gk, int n,
                                             const TileBuf<T> &tb, int nb,
                                             const LinVal<D, T, NC> &lv) {
    constexpr int K = D == 2 ? 9 : 27;
    __shared__ int wave_runs[NWAVE];
    int key = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) key = key * n + s.base[d];
    if (!valid) key = -1 - (int)(threadIdx.x & 63);
    const Seg sg = lane_runs(key);
    const RunMask<T> rm = run_mask<T>(sg.dist);
    // A lane without a particle takes part in the lane sums with zeros; its
    // registers are cleared here so that no stale value enters the sums.
    T wz[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) wz[i] = valid ? s.w[i][D - 1] : T(0);
    LinVal<D, T, NC> z;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        z.b[c] = valid ? lv.b[c] : T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) z.L[c][d] = valid && c < NL ? lv.L[c][d] : T(0);
    }

    // ---- staged path
    const bool tail = valid && sg.tail;
    const unsigned long long tmask = __ballot(tail);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wave_runs[wid] = __popcll(tmask);
    __syncthreads();
    int RW = 0;
#pragma unroll
    for (int w = 0; w < NWAVE; ++w) RW = max(RW, wave_runs[w]);
    int hcells = 1, nodes = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        hcells *= ext[d] + 2;
        nodes *= ext[d];
    }
    const size_t RC = (size_t)NWAVE * RW;
    const size_t stage_bytes = RC * K * 4 * sizeof(T);
    const int nflag_words = (nodes + 3) / 4;
    const bool staged = !tb.store && stage_bytes + (RC + hcells + nflag_words) * sizeof(int) <=
                                         (size_t)tb.tn * 4 * sizeof(T);
    if (staged) {
        T *stage = tile;                                                 // [RC][K][4]
        int *nxt = reinterpret_cast<int *>(reinterpret_cast<char *>(tile) + stage_bytes);
        int *head = nxt + RC;                                            // padded cells
        int *nflag_w = head + hcells;                                    // node touched
        unsigned char *nflag = reinterpret_cast<unsigned char *>(nflag_w);
        for (int i = threadIdx.x; i < hcells; i += blockDim.x) head[i] = -1;
        for (int i = threadIdx.x; i < nflag_words; i += blockDim.x) nflag_w[i] = 0;
        __syncthreads();
        const int rid = wid * RW + (int)__builtin_amdgcn_mbcnt_hi(
                                        (unsigned)(tmask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)tmask, 0u));
        if (tail) {
            int ci = 0, nl = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                ci = ci * (ext[d] + 2) + s.base[d] - lo[d] + 2;
                nl = nl * ext[d] + s.base[d] - lo[d];
            }
            nxt[rid] = atomicExch(head + ci, rid);
            // the run's 3^D nodes (the same byte may be set by several lanes)
#pragma unroll
            for (int o = 0; o < K; ++o) {
                int off[D], dl = 0;
                unflat<D>(o, off);
#pragma unroll
                for (int d = 0; d < D; ++d) dl = dl * ext[d] + off[d];
                nflag[nl + dl] = 1;
            }
        }
        run_offsets<D, T, NC, NL, U>(s, wz, z, rm, [&](const int *) { return valid; },
                                     [&](int o, const int *, const T v[NC]) {
            if (tail) {
                T r[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) r[c] = c < NC ? v[c] : T(0);
                st4(stage + ((size_t)rid * K + o) * 4, r);
            }
        });
        __syncthreads();
        for (int i = threadIdx.x; i < nodes; i += blockDim.x) {
            if (!nflag[i]) continue;
            int rem = i, c[D];
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                c[d] = rem % ext[d];
                rem /= ext[d];
            }
            int node = 0, ci0 = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                node = node * n + lo[d] + c[d];
                ci0 = ci0 * (ext[d] + 2) + c[d] + 2;
            }
            // Offsets in groups of 9: the heads of the group's cells, then
            // (unconditionally, from run 0 for an empty cell, masked) their
            // first runs' totals and list successors, so each step is one
            // round of independent LDS loads; longer lists are walked last.
            T acc[NC];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) acc[cc] = T(0);
#pragma unroll
            for (int g = 0; g < K; g += 9) {
                int h[9], h2[9];
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    int off[D], dl = 0;
                    unflat<D>(g + j, off);
#pragma unroll
                    for (int d = 0; d < D; ++d) dl = dl * (ext[d] + 2) + off[d];
                    h[j] = head[ci0 - dl];
                }
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    const int r = max(h[j], 0);
                    T v4[4];
                    ld4(stage + ((size_t)r * K + g + j) * 4, v4);
                    const int r2 = nxt[r];
                    h2[j] = h[j] >= 0 ? r2 : -1;
#pragma unroll
                    for (int cc = 0; cc < NC; ++cc) acc[cc] += h[j] >= 0 ? v4[cc] : T(0);
                }
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    for (int r = h2[j]; r >= 0; r = nxt[r]) {
                        T v4[4];
                        ld4(stage + ((size_t)r * K + g + j) * 4, v4);
#pragma unroll
                        for (int cc = 0; cc < NC; ++cc) acc[cc] += v4[cc];
                    }
                }
            }
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
                if (acc[cc] != T(0)) gadd(gk + (size_t)node * 4 + cc, acc[cc]);
        }
        __syncthreads();
        return;
    }

    // ---- dense tile path
    int plane = 1;
#pragma unroll
    for (int d = 1; d < D; ++d) plane *= ext[d];
    const bool direct = plane > tb.tn;
    const int sx = direct ? ext[0] : min(ext[0], tb.tn / plane);
    for (int x0 = 0; x0 < ext[0]; x0 += sx) {
        const int nx = min(sx, ext[0] - x0), snodes = nx * plane;
        if (!direct) {
            for (int i = threadIdx.x; i < snodes * 4; i += blockDim.x) tile[i] = T(0);
            __syncthreads();
        }
        run_offsets<D, T, NC, NL, U>(s, wz, z, rm, [&](const int *off) {
            const int l0 = s.base[0] - lo[0] + off[0] - x0;
            return valid && (direct || (l0 >= 0 && l0 < nx));
        }, [&](int, const int *off, const T v[NC]) {
            const int l0 = s.base[0] - lo[0] + off[0] - x0;
            if (!(tail && (direct || (l0 >= 0 && l0 < nx)))) return;
            if (direct) {
                int node = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) node = node * n + s.base[d] + off[d];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    if (v[c] != T(0)) gadd(gk + (size_t)node * 4 + c, v[c]);
            } else {
                int li = l0;
#pragma unroll
                for (int d = 1; d < D; ++d) li = li * ext[d] + s.base[d] - lo[d] + off[d];
#pragma unroll
                for (int c = 0; c < NC; ++c) atomicAdd(tile + li * 4 + c, v[c]);
            }
        });
        if (direct) continue;
        __syncthreads();
        if (tb.store && nx == ext[0] && store_tile<D, T>(tile, snodes, lo, ext, tb, nb, gk, n))
            break;
        for (int i = threadIdx.x; i < snodes; i += blockDim.x) {
            int rem = i, c[D];
#pragma unroll
            for (int d = D - 1; d >= 1; --d) {
                c[d] = lo[d] + rem % ext[d];
                rem /= ext[d];
            }
            c[0] = lo[0] + x0 + rem;
            int node = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) node = node * n + c[d];
#pragma unroll
            for (int cc = 0; cc < NC; ++cc) {
                const T r = tile[i * 4 + cc];
                if (r != T(0)) gadd(gk + (size_t)node * 4 + cc, r);
            }
        }
        __syncthreads();
    }
}

// Sensor sums of one workgroup's particles (3D): per group sum of x (3 per
// group), then sum x (3), sum v (3).  acc: LDS [3*Gs + 6], committed to the
// robot's sens row with one global atomic per nonzero slot.
template <typename T>
__device__ __forceinline__ void sense_block(bool valid, int g, const T nx[3], const T nv[3], int Gs,
                                            T *acc, T *sens_k) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T a = wave_sum(valid ? nx[d] : T(0));
        const T b = wave_sum(valid ? nv[d] : T(0));
        if (lane == 0) {
            atomicAdd(acc + 3 * Gs + d, a);
            atomicAdd(acc + 3 * Gs + 3 + d, b);
        }
    }
    const T xv[3] = {nx[0], nx[1], nx[2]};
    keyed_wave_add<T, 3>(valid, g, xv, acc, 3);
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * Gs + 6; i += blockDim.x)
        if (acc[i] != T(0)) gadd(sens_k + i, acc[i]);
}

// ---------------------------------------------------------------- controller state
template <typename T> struct Ctrl {
    int P, Gs, n_sin, run_period, I, H, O, act;
    T scale, max_speed;
    const T *inv_count, *targets, *W1, *b1, *W2, *b2;
    const int32_t *np, *hz;   // uneven batches (see MP); nullptr: even
    int per_robot_weights;    // robot k's W1, b1, W2, b2 at k (H I), k H, k (O H), k O
    int t;                    // the step (adjoint launches)
    // robot k's weights
    __device__ __forceinline__ const T *w1(int k) const { return W1 + (per_robot_weights ? (size_t)k * H * I : 0); }
    __device__ __forceinline__ const T *bb1(int k) const { return b1 + (per_robot_weights ? (size_t)k * H : 0); }
    __device__ __forceinline__ const T *w2(int k) const { return W2 + (per_robot_weights ? (size_t)k * O * H : 0); }
    __device__ __forceinline__ const T *bb2(int k) const { return b2 + (per_robot_weights ? (size_t)k * O : 0); }
};

// The controller of step t for robot k, run by every thread of a launch of
// nt = blockDim.x threads (a multiple of 64): the network input from the
// sensor sums (taichi_env.py:149-183 order for the MPM quadruped), hidden and
// output values (nn.py:130-150).  Layer 1 takes G = 4 threads per hidden unit
// when nt >= 4 H (partial sums over every G-th input, combined by lane
// shuffles), layer 2 one wavefront per output.  The outputs land in `so`
// (LDS, O values); with `rows` the center, COM velocity, input, hidden and
// output rows of step t are written too.  si, shd: LDS for I and H values.
// For a robot that takes step t (t < hz[k]); every thread must call it.
template <typename T>
__device__ __forceinline__ void ctrl_step(const Ctrl<T> &c, int k, int t, const T *sens, T *center,
                                          T *vcom, T *inp, T *hid, T *out, bool rows, T *si,
                                          T *shd, T *so) {
    const int tid = threadIdx.x, nt = blockDim.x, NS = 3 * c.Gs + 6;
    const T *sk = sens + (size_t)k * NS;
    const T invP = T(1) / T(c.np ? min(c.np[k], c.P) : c.P);
    if (rows && tid < 3) {
        center[k * 3 + tid] = sk[3 * c.Gs + tid] * invP;
        vcom[k * 3 + tid] = sk[3 * c.Gs + 3 + tid] * invP;
    }
    const T *W1 = c.w1(k), *b1 = c.bb1(k), *W2 = c.w2(k), *b2 = c.bb2(k);
    const T *tg = c.targets + (size_t)k * 5;
    for (int j = tid; j < c.I; j += nt) {
        T val;
        if (j < c.n_sin) {
            val = T(::sin(2.0 * M_PI / c.run_period * (double)t + 2.0 * M_PI / c.n_sin * (double)j));
        } else if (j < c.n_sin + 3) {
            val = sk[3 * c.Gs + 3 + (j - c.n_sin)] * invP;
        } else if (j < c.n_sin + 3 + 3 * c.Gs) {
            const int q = j - c.n_sin - 3, g = q / 3, d = q - 3 * g;
            val = (sk[q] * c.inv_count[(size_t)k * c.Gs + g] - sk[3 * c.Gs + d] * invP) / c.scale;
        } else if (j < c.n_sin + 6 + 3 * c.Gs) {
            val = tg[j - c.n_sin - 3 - 3 * c.Gs] / c.max_speed;
        } else {
            val = tg[3];
        }
        si[j] = val;
        if (rows) inp[(size_t)k * c.I + j] = val;
    }
    __syncthreads();
    const int G = nt >= 4 * c.H ? 4 : 1;
    for (int h0 = 0; h0 < c.H; h0 += nt / G) {
        const int h = h0 + tid / G, part = tid % G;
        T z = T(0);
        if (h < c.H) {
            const T *w = W1 + (size_t)h * c.I;
            for (int i = part; i < c.I; i += G) z += w[i] * si[i];
        }
        if (G == 4) {
            z += __shfl_xor(z, 1, 64);
            z += __shfl_xor(z, 2, 64);
        }
        if (h < c.H && part == 0) {
            T y, dy;
            activate<T>(c.act, z + b1[h], y, dy);
            shd[h] = y;
            if (rows) hid[(size_t)k * c.H + h] = y;
        }
    }
    __syncthreads();
    const int lane = tid & 63;
    for (int o = tid >> 6; o < c.O; o += nt >> 6) {
        T z = T(0);
        for (int h = lane; h < c.H; h += 64) z += W2[(size_t)o * c.H + h] * shd[h];
        z = wave_sum(z);
        if (lane == 0) {
            T y, dy;
            activate<T>(c.act, z + b2[o], y, dy);
            so[o] = y;
            if (rows) out[(size_t)k * c.O + o] = y;
        }
    }
    __syncthreads();
}

// Controller rows of one step (CTRL launches of P2G): row pointers at step t.
template <typename T> struct CtrlRows {
    const T *sens;
    T *center, *vcom, *inp, *hid, *out;
};

// ---------------------------------------------------------------- P2G
// Particle range of this workgroup: window blockIdx.x of the window table,
// or the blockIdx.x-th chunk of PB particles without one.
struct Window {
    int q0, q1;
};
__device__ __forceinline__ Window window_of(const int32_t *win, int P) {
    if (!win) return {(int)blockIdx.x * PB, min((int)blockIdx.x * PB + PB, P)};
    return {win[2 * blockIdx.x], min(win[2 * blockIdx.x + 1], P)};
}

// CTRL: the step's controller runs at the start of every workgroup (after the
// particle loads are issued, so their latency overlaps it) instead of in a
// launch of its own; workgroup 0 of each robot writes the controller rows
// (and the center / COM velocity of the final state of a robot whose horizon
// ends at this step).  Otherwise the actuation comes from act.
template <int D, typename T, bool STORE_F, int U, bool CTRL = false>
__global__ __launch_bounds__(PB) void k_p2g_b(MP<T> p, PView<D, T> pv, const T *act,
                                              const int32_t *act_id, int n_act, T *F_out,
                                              T *grid, Blk bk, int32_t *err, TileBuf<T> tb,
                                              const int32_t *win, Ctrl<T> cc = {},
                                              CtrlRows<T> cr = {}) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);     // [tb.tn][4]
    __shared__ int red[2 * D * NWAVE];
    const int k = blockIdx.y;
    if (idle_at(p, k, p.t)) {
        if constexpr (CTRL) {
            if (p.t == cc.hz[k] && blockIdx.x == 0 && threadIdx.x < 3) {
                const T *sk = cr.sens + (size_t)k * (3 * cc.Gs + 6);
                const T invP = T(1) / T(cc.np ? min(cc.np[k], cc.P) : cc.P);
                cr.center[k * 3 + threadIdx.x] = sk[3 * cc.Gs + threadIdx.x] * invP;
                cr.vcom[k * 3 + threadIdx.x] = sk[3 * cc.Gs + 3 + threadIdx.x] * invP;
            }
        }
        return;
    }
    const Window wn = window_of(win, n_part(p, k, pv.P));
    if (threadIdx.x == 0 && wn.q1 - wn.q0 > PB) atomicOr(err, 2);
    const int q = wn.q0 + (int)threadIdx.x, P = pv.P;
    bool valid = q >= 0 && q < wn.q1;
    T mv[D];
    Mat<D, T> aff;
    Stencil<D, T> s;
#pragma unroll
    for (int d = 0; d < D; ++d) s.base[d] = 0;
    T xs[D], vs[D];
    Mat<D, T> Cm, Fm;
    int id = -1;
    if (valid) {
        pv.load(k, q, xs, vs, Cm, Fm);
        id = act_id[(size_t)k * P + q];
    }
    T a = T(0);
    if constexpr (CTRL) {
        T *si = tile, *shd = si + cc.I, *so = shd + cc.H;
        ctrl_step<T>(cc, k, p.t, cr.sens, cr.center, cr.vcom, cr.inp, cr.hid, cr.out,
                     blockIdx.x == 0, si, shd, so);
        if (valid && id >= 0) a = so[id];
    } else {
        if (valid && id >= 0) a = act[(size_t)k * n_act + id];
    }
    if (valid) {
        const Consti<D, T> c = constitutive<D, T>(Cm, Fm, a, p);
        if (STORE_F) store_mat<D, T>(F_out, k, q, P, c.Fn, false);
        aff = c.affine;
#pragma unroll
        for (int d = 0; d < D; ++d) mv[d] = p.p_mass * vs[d];
        s = stencil<D, T>(xs, p);
        if (!s.ok) {
            atomicOr(err, 1);
            valid = false;
        }
    }
    int lo[D], ext[D];
    if (!block_bbox<D>(valid, s.base, lo, ext, red)) return;
    T *gk = grid + (size_t)k * gsize<D>(p.n) * 4;
    // momentum of node base + off: W (mv + A (off - fx) dx), mass W m_p
    LinVal<D, T, 4> lv;
#pragma unroll
    for (int i = 0; i < D; ++i) {
        T b = mv[i];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            lv.L[i][j] = aff.a[i][j] * p.dx;
            b -= lv.L[i][j] * s.fx[j];
        }
        lv.b[i] = b;
    }
#pragma unroll
    for (int i = D; i < 3; ++i) lv.b[i] = T(0);
    lv.b[3] = p.p_mass;
    tile_scatter<D, T, 4, D, U>(valid, s, lo, ext, tile, gk, p.n, tb.robot(k, bk.nblk), bk.nb,
                                lv);
    mark_box<D>(bk, k, lo, ext);
}

// ---------------------------------------------------------------- grid
// One step's saved grid (segment step j): records by active-list position.
template <typename T> struct GridSave {
    T *g, *v;              // [B][cap][64][4] grid totals, velocities
    int32_t *list, *count; // [B][cap], [B]
    int cap;               // records per robot (blocks by active-list position)
    int32_t *over;         // set when a robot has more active blocks than cap (then
                           // nothing of that robot is saved); may be null
};

// Velocity of every node of the active blocks; with `clear`, the momentum /
// mass records are zeroed after use (forward); flags are always reset.  With
// sv.g set, the totals, velocities and the active list are saved too.
template <int D, typename T>
__global__ __launch_bounds__(PB) void k_grid_b(MP<T> p, T *grid, T *gvo, Blk bk, int clear,
                                               TileBuf<T> tb, GridSave<T> sv) {
    const int k = blockIdx.y, lane = threadIdx.x & 63;
    if (idle_at(p, k, p.t)) return;
    const TileBuf<T> tbk = tb.robot(k, bk.nblk);
    const int nw = gridDim.x * NWAVE;
    const int cnt = bk.count[k];
    const bool save = sv.g && cnt <= sv.cap;
    if (sv.g && blockIdx.x == 0 && threadIdx.x == 0) {
        if (save)
            sv.count[k] = cnt;
        else if (sv.over)
            atomicOr(sv.over, 1);
    }
    const int32_t *lst = bk.list + (size_t)k * bk.nblk;
    const size_t G = gsize<D>(p.n);
    T *gk = grid + (size_t)k * G * 4, *ok = gvo + (size_t)k * G * 4;
    for (int e = blockIdx.x * NWAVE + (threadIdx.x >> 6); e < cnt; e += nw) {
        const int bid = lst[e];
        bool bnd;
        int c[D];
        const int node = block_node<D, T>(bid, lane, bk.nb, p, bnd, c);
        T r[4], o[4] = {T(0), T(0), T(0), T(0)};
        ld4(gk + (size_t)node * 4, r);
        gather_tiles<D, T>(tbk, bid, c, r);
        const T inv_m = T(1) / (r[3] + T(1e-10));
#pragma unroll
        for (int d = 0; d < D; ++d) {
            T vo = inv_m * r[d];
            if (d == 1) vo += p.g_dt;
            o[d] = bnd ? T(0) : vo;
        }
        st4(ok + (size_t)node * 4, o);
        if (save) {
            const size_t si = (((size_t)k * sv.cap + e) * 64 + lane) * 4;
            st4(sv.g + si, r);
            st4(sv.v + si, o);
            if (lane == 0) sv.list[(size_t)k * sv.cap + e] = bid;
        }
        if (clear) {
            const T z[4] = {T(0), T(0), T(0), T(0)};
            st4(gk + (size_t)node * 4, z);
        } else {
            st4(gk + (size_t)node * 4, r);      // totals, read by the grid adjoint
        }
        if (lane == 0) {
            bk.flag[(size_t)k * bk.nblk + bid] = 0;
            tbk.cnt[bid] = 0;
        }
    }
}

// Restore a saved step: grid totals and velocities of its active blocks,
// and the active list itself, as the P2G + grid passes left them.  With
// vo_only only the velocities are restored (the replayed recompute's G2P
// reads nothing else); otherwise also the totals, the active list and
// count, which the adjoint kernels consume and clear.
template <int D, typename T>
__device__ __forceinline__ void grid_restore(const MP<T> &p, T *grid, T *gvo, const Blk &bk,
                                             const GridSave<T> &sv, int vo_only, int bx, int nbx) {
    const int k = blockIdx.y, lane = threadIdx.x & 63;
    const int nw = nbx * NWAVE;
    const int cnt = sv.count[k];
    const size_t G = gsize<D>(p.n);
    T *gk = grid + (size_t)k * G * 4, *ok = gvo + (size_t)k * G * 4;
    if (!vo_only && bx == 0 && threadIdx.x == 0) bk.count[k] = cnt;
    for (int e = bx * NWAVE + (threadIdx.x >> 6); e < cnt; e += nw) {
        const int bid = sv.list[(size_t)k * sv.cap + e];
        bool bnd;
        int c[D];
        const int node = block_node<D, T>(bid, lane, bk.nb, p, bnd, c);
        const size_t si = (((size_t)k * sv.cap + e) * 64 + lane) * 4;
        T o[4];
        ld4(sv.v + si, o);
        st4(ok + (size_t)node * 4, o);
        if (!vo_only) {
            T r[4];
            ld4(sv.g + si, r);
            st4(gk + (size_t)node * 4, r);
            if (lane == 0) bk.list[(size_t)k * bk.nblk + e] = bid;
        }
    }
}

template <int D, typename T>
__global__ __launch_bounds__(PB) void k_grid_restore(MP<T> p, T *grid, T *gvo, Blk bk,
                                                     GridSave<T> sv, int vo_only) {
    if (idle_at(p, blockIdx.y, p.t)) return;
    grid_restore<D, T>(p, grid, gvo, bk, sv, vo_only, blockIdx.x, gridDim.x);
}

// F of the next state of particle q of robot k, (I + dt C) F: the part of
// P2G that a replayed recompute (saved grid velocities + G2P) still needs.
template <int D, typename T>
__device__ __forceinline__ void f_update(const MP<T> &p, int P, const T *C, const T *F, T *F_out,
                                         int k, int q) {
    const size_t bm = (size_t)k * D * D * P;
    Mat<D, T> Cm, Fm;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            Cm.a[i][j] = C[bm + (size_t)(i * D + j) * P + q];
            Fm.a[i][j] = F[bm + (size_t)(i * D + j) * P + q];
        }
    store_mat<D, T>(F_out, k, q, P, f_next<D, T>(Cm, Fm, p), false);
}


// ---------------------------------------------------------------- G2P
template <typename T> struct SenseOut {
    T *sens;               // row t+1: [B][3*Gs + 6]
    const int32_t *group;  // [B][P]
    int Gs;
};

// FUPD: also the F update of each particle (f_update: C_in, F_in -> F_out),
// for the replayed recompute, which has no P2G to do it.
template <int D, typename T, bool SENSE, bool FUPD = false>
__global__ __launch_bounds__(PB) void k_g2p_b(MP<T> p, int P, const T *x, const T *gvo, T *x_out,
                                              T *v_out, T *C_out, int32_t *count,
                                              SenseOut<T> so, const T *C_in = nullptr,
                                              const T *F_in = nullptr, T *F_out = nullptr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *acc = reinterpret_cast<T *>(smem_raw);
    const int k = blockIdx.y, q = blockIdx.x * PB + threadIdx.x;
    if (idle_at(p, k, p.t)) return;
    if (SENSE) {
        for (int i = threadIdx.x; i < 3 * so.Gs + 6; i += PB) acc[i] = T(0);
        __syncthreads();
    }
    // the grid pass that read the active list has finished: empty it
    if (blockIdx.x == 0 && threadIdx.x == 0) count[k] = 0;
    const bool valid = q < n_part(p, k, P);
    if constexpr (FUPD)
        if (valid) f_update<D, T>(p, P, C_in, F_in, F_out, k, q);
    T nx[D], nv[D];
#pragma unroll
    for (int d = 0; d < D; ++d) nx[d] = nv[d] = T(0);
    if (valid) {
        const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
        T xs[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xs[d] = x[bv + (size_t)d * P + q];
        const Stencil<D, T> s = stencil<D, T>(xs, p);
        T nC[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) nC[i][j] = T(0);
        if (s.ok) {
            const T *g = gvo + (size_t)k * gsize<D>(p.n) * 4;
            constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
            for (int o = 0; o < K; ++o) {
                int off[D];
                unflat<D>(o, off);
                T W = T(1), dp[D];
                int node = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    W *= s.w[off[d]][d];
                    dp[d] = T(off[d]) - s.fx[d];
                    node = node * p.n + s.base[d] + off[d];
                }
                T gv4[4];
                ld4(g + (size_t)node * 4, gv4);
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    nv[i] += W * gv4[i];
#pragma unroll
                    for (int j = 0; j < D; ++j) nC[i][j] += T(4) * p.inv_dx * W * gv4[i] * dp[j];
                }
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            nx[d] = xs[d] + p.dt * nv[d];
            v_out[bv + (size_t)d * P + q] = nv[d];
            x_out[bv + (size_t)d * P + q] = nx[d];
        }
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) C_out[bm + (size_t)(i * D + j) * P + q] = nC[i][j];
    }
    if constexpr (SENSE && D == 3) {
        const int g = valid ? so.group[(size_t)k * P + q] : 0;
        sense_block<T>(valid, g, nx, nv, so.Gs, acc, so.sens + (size_t)k * (3 * so.Gs + 6));
    }
}

// Sensor sums of a stored state (the first step of a rollout).
template <typename T>
__global__ __launch_bounds__(PB) void k_sense(int P, const int32_t *np, const T *x, const T *v,
                                              SenseOut<T> so) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *acc = reinterpret_cast<T *>(smem_raw);
    const int k = blockIdx.y, q = blockIdx.x * PB + threadIdx.x;
    for (int i = threadIdx.x; i < 3 * so.Gs + 6; i += PB) acc[i] = T(0);
    __syncthreads();
    const bool valid = q < (np ? min(np[k], P) : P);
    T nx[3] = {T(0), T(0), T(0)}, nv[3] = {T(0), T(0), T(0)};
    int g = 0;
    if (valid) {
        const size_t bv = (size_t)k * 3 * P;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            nx[d] = x[bv + (size_t)d * P + q];
            nv[d] = v[bv + (size_t)d * P + q];
        }
        g = so.group[(size_t)k * P + q];
    }
    sense_block<T>(valid, g, nx, nv, so.Gs, acc, so.sens + (size_t)k * (3 * so.Gs + 6));
}

// ---------------------------------------------------------------- controller
// COM and COM velocity of the final state (step t = horizon) from its sensor
// sums, one 64-thread block per robot; the steps before it compute theirs
// with the controller inside P2G (ctrl_step).
template <typename T>
__global__ __launch_bounds__(64) void k_center(Ctrl<T> c, int t, const T *sens, T *center,
                                               T *vcom) {
    const int k = blockIdx.x, tid = threadIdx.x;
    // robot k: state t exists for t <= hz[k] (its last one written in P2G)
    if (c.hz && t != c.hz[k]) return;
    const T *sk = sens + (size_t)k * (3 * c.Gs + 6);
    const T invP = T(1) / T(c.np ? min(c.np[k], c.P) : c.P);
    if (tid < 3) {
        center[k * 3 + tid] = sk[3 * c.Gs + tid] * invP;
        vcom[k * 3 + tid] = sk[3 * c.Gs + 3 + tid] * invP;
    }
}

// Adjoint of the controller (ctrl_step) at step t for robot k, by the nt threads of
// a workgroup (each entry computed by one thread, in the same order for any
// nt): consumes (and zeroes) act_b, writes in_b row t and the pre-activation
// adjoints of the step (zh: hidden [B][H], zo: output [B][O]); k_wgrad_time
// turns those rows into the weight gradient.  LDS: (I + 3H + O) values.
template <typename T>
__device__ __forceinline__ void ctrl_adj(const Ctrl<T> &c, int k, int nt, const T *inp, T *act_b,
                                         T *in_b, T *zh, T *zo) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *si = reinterpret_cast<T *>(smem_raw);
    T *shd = si + c.I;
    T *sd1 = shd + c.H;
    T *sz1b = sd1 + c.H;
    T *sab = sz1b + c.H;
    const int tid = threadIdx.x;
    const T *W1 = c.w1(k), *b1 = c.bb1(k), *W2 = c.w2(k), *b2 = c.bb2(k);
    for (int i = tid; i < c.I; i += nt) si[i] = inp[(size_t)k * c.I + i];
    __syncthreads();
    for (int h = tid; h < c.H; h += nt) {
        T z = b1[h];
        const T *w = W1 + (size_t)h * c.I;
        for (int i = 0; i < c.I; ++i) z += w[i] * si[i];
        T y, dy;
        activate<T>(c.act, z, y, dy);
        shd[h] = y;
        sd1[h] = dy;
    }
    __syncthreads();
    for (int o = tid; o < c.O; o += nt) {
        T z = b2[o];
        const T *w = W2 + (size_t)o * c.H;
        for (int h = 0; h < c.H; ++h) z += w[h] * shd[h];
        T y, dy;
        activate<T>(c.act, z, y, dy);
        T *ab = act_b + (size_t)k * c.O + o;
        sab[o] = *ab * dy;
        *ab = T(0);
    }
    __syncthreads();
    for (int h = tid; h < c.H; h += nt) {
        T hb = T(0);
        for (int o = 0; o < c.O; ++o) hb += W2[(size_t)o * c.H + h] * sab[o];
        sz1b[h] = hb * sd1[h];
    }
    __syncthreads();
    for (int i = tid; i < c.I; i += nt) {
        T ib = T(0);
        for (int h = 0; h < c.H; ++h) ib += W1[(size_t)h * c.I + i] * sz1b[h];
        in_b[(size_t)k * c.I + i] = ib;
    }
    for (int h = tid; h < c.H; h += nt) zh[(size_t)k * c.H + h] = sz1b[h];
    for (int o = tid; o < c.O; o += nt) zo[(size_t)k * c.O + o] = sab[o];
}

// Block per robot: ctrl_adj on 64 threads.
template <typename T>
__global__ __launch_bounds__(64) void k_ctrl_adj(Ctrl<T> c, const T *inp, T *act_b, T *in_b,
                                                 T *zh, T *zo) {
    if (c.hz && c.t >= c.hz[blockIdx.x]) return;
    ctrl_adj<T>(c, blockIdx.x, 64, inp, act_b, in_b, zh, zo);
}

// Weight gradient of n controller steps, added once to the robot's slice
// [W1 (H*I) | b1 (H) | W2 (O*H) | b2 (O)] of wg: dW1 = sum_t zh_t inp_t^T,
// db1 = sum_t zh_t, dW2 = sum_t zo_t hid_t^T, db2 = sum_t zo_t (nn.py:130-150
// chain rule).  Row pointers start at the first step; rows are [B][.] per
// step.  One thread per weight entry, grid (entries / 256, B); steps are
// summed from the last to the first, the order the per-step adjoints run in.
template <typename T>
__global__ __launch_bounds__(256) void k_wgrad_time(int B, int I, int H, int O, int n,
                                                    const T *inp, const T *hid, const T *zh,
                                                    const T *zo, T *wg, const int32_t *hz,
                                                    int t0) {
    const int k = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;
    const int nW1 = H * I, nb1 = nW1 + H, nW2 = nb1 + O * H, nw = nW2 + O;
    if (e >= nw) return;
    if (hz) n = max(0, min(n, hz[k] - t0));   // robot k's steps of the segment
    T acc = T(0);
    if (e < nW1) {
        const int h = e / I, i = e - h * I;
        for (int j = n - 1; j >= 0; --j)
            acc += zh[((size_t)j * B + k) * H + h] * inp[((size_t)j * B + k) * I + i];
    } else if (e < nb1) {
        const int h = e - nW1;
        for (int j = n - 1; j >= 0; --j) acc += zh[((size_t)j * B + k) * H + h];
    } else if (e < nW2) {
        const int r = e - nb1, o = r / H, h = r - o * H;
        for (int j = n - 1; j >= 0; --j)
            acc += zo[((size_t)j * B + k) * O + o] * hid[((size_t)j * B + k) * H + h];
    } else {
        const int o = e - nW2;
        for (int j = n - 1; j >= 0; --j) acc += zo[((size_t)j * B + k) * O + o];
    }
    wg[(size_t)k * nw + e] += acc;
}

// One launch for two independent pieces of the backward: the controller
// adjoint of step t (workgroup nr of each robot, after P2G's adjoint of step
// t has written act_b) and the grid restore of step t - 1 (workgroups
// 0 .. nr-1), which the adjoint of step t - 1 reads next.
template <int D, typename T>
__global__ __launch_bounds__(PB) void k_restore_ctrl_adj(MP<T> p, T *grid, T *gvo, Blk bk,
                                                         GridSave<T> sv, int nr, Ctrl<T> c,
                                                         const T *inp, T *act_b, T *in_b, T *zh,
                                                         T *zo) {
    if ((int)blockIdx.x < nr) {
        if (!idle_at(p, blockIdx.y, p.t - 1))
            grid_restore<D, T>(p, grid, gvo, bk, sv, 0, blockIdx.x, nr);
    } else if (!idle_at(p, blockIdx.y, p.t)) {
        ctrl_adj<T>(c, blockIdx.y, PB, inp, act_b, in_b, zh, zo);
    }
}

// ---------------------------------------------------------------- adjoints
template <typename T> struct SenseAdj {
    const T *in_b;         // row t+1: [B][I]; null: no sensor terms
    const T *cg, *vg;      // rows t+1: [B][3]
    const int32_t *group;  // [B][P]
    const T *inv_count;    // [B][Gs]
    int Gs, n_sin, I;
    T scale;
};

// G2P adjoint with the sensor / COM adjoint of state t+1 folded into the
// incoming (x, v) adjoints: writes xb (state t), scatters dL/d grid_vo.
template <int D, typename T, int U>
__global__ __launch_bounds__(PB) void k_g2p_adj_b(MP<T> p, int P, const T *x, const T *gvo,
                                                  const T *xb1, const T *vb1, const T *Cb1, T *xb,
                                                  T *gvo_b, SenseAdj<T> sa, Blk bk, TileBuf<T> tb,
                                                  const int32_t *win) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *tile = reinterpret_cast<T *>(smem_raw);     // [tb.tn][4]
    __shared__ int red[2 * D * NWAVE];
    if (idle_at(p, blockIdx.y, p.t)) return;
    const Window wn = window_of(win, n_part(p, blockIdx.y, P));
    const int k = blockIdx.y, q = wn.q0 + (int)threadIdx.x;
    const bool valid = q >= 0 && q < wn.q1;
    const size_t bv = (size_t)k * D * P, bm = (size_t)k * D * D * P;
    T xs[D], xe[D], vnb[D], Cb[D][D];
    Stencil<D, T> s;
#pragma unroll
    for (int d = 0; d < D; ++d) s.base[d] = 0;
    bool inside = false;
    if (valid) {
        T ve[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xs[d] = x[bv + (size_t)d * P + q];
            xe[d] = xb1[bv + (size_t)d * P + q];
            ve[d] = vb1[bv + (size_t)d * P + q];
        }
        if constexpr (D == 3) {
            if (sa.in_b) {
                const T *ib = sa.in_b + (size_t)k * sa.I;
                const int g = sa.group[(size_t)k * P + q];
                const T invP = T(1) / T(n_part(p, k, P));
                const T icg = sa.inv_count[(size_t)k * sa.Gs + g];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T S = T(0);
                    for (int h = 0; h < sa.Gs; ++h) S += ib[sa.n_sin + 3 + 3 * h + d];
                    xe[d] += (ib[sa.n_sin + 3 + 3 * g + d] * icg - S * invP) / sa.scale +
                             sa.cg[k * 3 + d] * invP;
                    ve[d] += (ib[sa.n_sin + d] + sa.vg[k * 3 + d]) * invP;
                }
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) vnb[d] = ve[d] + p.dt * xe[d];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) Cb[i][j] = Cb1[bm + (size_t)(i * D + j) * P + q];
        s = stencil<D, T>(xs, p);
        inside = s.ok;
    }
    // gather: stencil-weight derivatives against grid_vo
    T fxb[D];
#pragma unroll
    for (int d = 0; d < D; ++d) fxb[d] = T(0);
    const size_t G = gsize<D>(p.n);
    const T c4 = T(4) * p.inv_dx;
    if (inside) {
        const T *g = gvo + (size_t)k * G * 4;
        constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
        for (int o = 0; o < K; ++o) {
            int off[D];
            unflat<D>(o, off);
            T W = T(1), dp[D];
            int node = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                W *= s.w[off[d]][d];
                dp[d] = T(off[d]) - s.fx[d];
                node = node * p.n + s.base[d] + off[d];
            }
            T vo4[4];
            ld4(g + (size_t)node * 4, vo4);
            T Wb = T(0);
#pragma unroll
            for (int i = 0; i < D; ++i) {
                T acc = T(0);
#pragma unroll
                for (int j = 0; j < D; ++j) acc += Cb[i][j] * dp[j];
                Wb += vo4[i] * (vnb[i] + c4 * acc);
            }
#pragma unroll
            for (int j = 0; j < D; ++j) {
                T CtV = T(0);
#pragma unroll
                for (int i = 0; i < D; ++i) CtV += Cb[i][j] * vo4[i];
                fxb[j] -= c4 * W * CtV;
            }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                T dW = s.dw[off[d]][d];
#pragma unroll
                for (int e = 0; e < D; ++e)
                    if (e != d) dW *= s.w[off[e]][e];
                fxb[d] += Wb * dW;
            }
        }
    }
    // scatter: dL/d grid_vo += W (vnb + 4/dx Cb dpos)
    int lo[D], ext[D];
    if (block_bbox<D>(inside, s.base, lo, ext, red)) {
        // W (vnb + 4/dx Cb (off - fx))
        LinVal<D, T, D> lv;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            T b = vnb[i];
#pragma unroll
            for (int j = 0; j < D; ++j) {
                lv.L[i][j] = c4 * Cb[i][j];
                b -= lv.L[i][j] * s.fx[j];
            }
            lv.b[i] = b;
        }
        tile_scatter<D, T, D, D, U>(inside, s, lo, ext, tile, gvo_b + (size_t)k * G * 4, p.n,
                                    tb.robot(k, bk.nblk), bk.nb, lv);
    }
    if (valid) {
#pragma unroll
        for (int d = 0; d < D; ++d) xb[bv + (size_t)d * P + q] = xe[d] + p.inv_dx * fxb[d];
    }
}

// Grid adjoint over the active blocks: dL/d(momentum, mass) from dL/d vo;
// clears the grid and vo-adjoint records it consumed.
template <int D, typename T>
__global__ __launch_bounds__(PB) void k_grid_adj_b(MP<T> p, T *grid, T *gvo_b, T *grid_b, Blk bk,
                                                   TileBuf<T> tb) {
    const int k = blockIdx.y, lane = threadIdx.x & 63;
    if (idle_at(p, k, p.t)) return;
    const TileBuf<T> tbk = tb.robot(k, bk.nblk);
    const int nw = gridDim.x * NWAVE;
    const int cnt = bk.count[k];
    const int32_t *lst = bk.list + (size_t)k * bk.nblk;
    const size_t G = gsize<D>(p.n);
    T *gk = grid + (size_t)k * G * 4, *obk = gvo_b + (size_t)k * G * 4;
    T *ibk = grid_b + (size_t)k * G * 4;
    const T z[4] = {T(0), T(0), T(0), T(0)};
    for (int e = blockIdx.x * NWAVE + (threadIdx.x >> 6); e < cnt; e += nw) {
        bool bnd;
        int c[D];
        const int bid = lst[e];
        const int node = block_node<D, T>(bid, lane, bk.nb, p, bnd, c);
        T r[4], ob[4], o[4] = {T(0), T(0), T(0), T(0)};
        ld4(gk + (size_t)node * 4, r);
        ld4(obk + (size_t)node * 4, ob);
        gather_tiles<D, T>(tbk, bid, c, ob);
        const T m = r[3] + T(1e-10);
        T mb = T(0);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T b = bnd ? T(0) : ob[d];
            o[d] = b / m;
            mb -= b * r[d] / (m * m);
        }
        o[3] = mb;
        st4(ibk + (size_t)node * 4, o);
        st4(gk + (size_t)node * 4, z);
        st4(obk + (size_t)node * 4, z);
        if (lane == 0) tbk.cnt[bid] = 0;
    }
}

// P2G adjoint: dL/d(momentum, mass) and dL/dF[t+1] -> xb (+=), vb, Cb, Fb
// (written), act_b (wave-aggregated atomics).  Empties the active list.
template <int D, typename T>
__global__ __launch_bounds__(PB) void k_p2g_adj_b(MP<T> p, PView<D, T> pv, const T *act,
                                                  const int32_t *act_id, int n_act,
                                                  const T *grid_b, const T *Fb1, T *xb, T *vb,
                                                  T *Cb, T *Fb, T *act_b, int32_t *count) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *sab = reinterpret_cast<T *>(smem_raw);     // [n_act]
    const int k = blockIdx.y, q = blockIdx.x * PB + threadIdx.x, P = pv.P;
    if (idle_at(p, k, p.t)) return;
    for (int i = threadIdx.x; i < n_act; i += PB) sab[i] = T(0);
    __syncthreads();
    // the grid adjoint pass that read the active list has finished
    if (blockIdx.x == 0 && threadIdx.x == 0) count[k] = 0;
    const bool valid = q < n_part(p, k, P);
    int id = -1;
    T abar[1] = {T(0)};
    if (valid) {
        T xs[D], vs[D];
        Mat<D, T> Cm, Fm;
        pv.load(k, q, xs, vs, Cm, Fm);
        id = act_id[(size_t)k * P + q];
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
                Fnb.a[i][j] = Fb1[bm + (size_t)(i * D + j) * P + q];
            }
        T vacc[D], fxb[D];
#pragma unroll
        for (int d = 0; d < D; ++d) vacc[d] = fxb[d] = T(0);
        if (s.ok) {
            const T *gb = grid_b + (size_t)k * gsize<D>(p.n) * 4;
            constexpr int K = D == 2 ? 9 : 27;
#pragma unroll
            for (int o = 0; o < K; ++o) {
                int off[D];
                unflat<D>(o, off);
                T W = T(1), dpos[D];
                int node = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    W *= s.w[off[d]][d];
                    dpos[d] = (T(off[d]) - s.fx[d]) * p.dx;
                    node = node * p.n + s.base[d] + off[d];
                }
                T g4[4];
                ld4(gb + (size_t)node * 4, g4);
                T Wb = g4[3] * p.p_mass;
#pragma unroll
                for (int i = 0; i < D; ++i) {
                    T cv = p.p_mass * vs[i];
#pragma unroll
                    for (int j = 0; j < D; ++j) cv += c.affine.a[i][j] * dpos[j];
                    Wb += g4[i] * cv;
                    vacc[i] += W * p.p_mass * g4[i];
#pragma unroll
                    for (int j = 0; j < D; ++j) affb.a[i][j] += W * g4[i] * dpos[j];
                }
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    T atg = T(0);
#pragma unroll
                    for (int i = 0; i < D; ++i) atg += c.affine.a[i][j] * g4[i];
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
        abar[0] = consti_adj<D, T>(p, c, Cm, Fm, affb, Fnb, Fbm, Cbm);
        store_mat<D, T>(Fb, k, q, P, Fbm, false);
        store_mat<D, T>(Cb, k, q, P, Cbm, false);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            vb[bv + (size_t)d * P + q] = vacc[d];
            xb[bv + (size_t)d * P + q] += p.inv_dx * fxb[d];
        }
    }
    keyed_wave_add<T, 1>(valid && id >= 0, id, abar, sab, 1);
    __syncthreads();
    for (int i = threadIdx.x; i < n_act; i += PB)
        if (sab[i] != T(0)) gadd(act_b + (size_t)k * n_act + i, sab[i]);
}

template <typename T>
__global__ void k_copy(const T *src, T *dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace dl

// ================================================================== C ABI
using namespace dl;

// Unroll factor of the scatter kernels' stencil-offset loop (registers vs
// instruction-level parallelism), chosen at run time: 1, 3, 9 or 27.
#define DL_UNROLL_DISPATCH(U_, ...)                 \
    do {                                            \
        switch (unroll) {                           \
        case 1: { constexpr int U_ = 1; __VA_ARGS__; } break;   \
        case 9: { constexpr int U_ = 9; __VA_ARGS__; } break;   \
        case 27: { constexpr int U_ = 27; __VA_ARGS__; } break; \
        default: { constexpr int U_ = 3; __VA_ARGS__; } break;  \
        }                                           \
    } while (0)

namespace {

int check_batch(const dl_mpm_batch *b) {
    DL_CHECK_ARG(b != nullptr, "MPM batch descriptor is NULL");
    const dl_mpm_params &c = b->params;
    DL_CHECK_ARG(c.dim == 2 || c.dim == 3, "MPM batch: dim must be 2 or 3");
    const int E = c.dim == 3 ? 4 : 8;
    DL_CHECK_ARG(c.n_grid >= 8 && c.n_grid <= 1024 && c.n_grid % E == 0,
                 "MPM batch: n_grid must be a multiple of %d in [8, 1024]", E);
    DL_CHECK_ARG(c.act_axis >= 0 && c.act_axis < c.dim, "MPM batch: bad actuation axis");
    DL_CHECK_ARG(b->dtype == DL_F32 || b->dtype == DL_F64, "MPM batch: bad dtype");
    DL_CHECK_ARG(b->tile_nodes >= 64 && b->tile_nodes <= DL_MPM_TILE_MAX &&
                     b->tile_nodes * 4 * (b->dtype == DL_F64 ? 8 : 4) <= 128 * 1024,
                 "MPM batch: tile_nodes %d outside [64, %d] or over 128 KiB of LDS",
                 b->tile_nodes, (int)DL_MPM_TILE_MAX);
    DL_CHECK_ARG(b->scatter == DL_MPM_SCATTER_ATOMIC || b->scatter == DL_MPM_SCATTER_STORE,
                 "MPM batch: bad scatter mode %d", b->scatter);
    DL_CHECK_ARG(b->B > 0 && b->P > 0 && b->n_act > 0 && b->horizon > 0, "MPM batch: bad sizes");
    DL_CHECK_ARG(!(b->n_part || b->horizons || b->per_robot_weights) ||
                     (!b->windows && !b->hgrid && b->restore_from == 0),
                 "MPM batch: uneven batches (n_part, horizons, per-robot weights) run with plain "
                 "windows and without the horizon save");
    DL_CHECK_ARG(b->per_robot_weights == 0 || b->per_robot_weights == 1,
                 "MPM batch: per_robot_weights must be 0 or 1");
    DL_CHECK_ARG(!b->windows || (b->n_windows >= (b->P + PB - 1) / PB && b->n_windows <= b->P),
                 "MPM batch: %d windows cannot partition %d particles in windows of at most %d",
                 b->n_windows, b->P, PB);
    DL_CHECK_ARG(b->act_id && b->grid && b->grid_vo && b->blk_flag && b->blk_list &&
                     b->blk_count && b->err && b->act_out && b->tiles && b->tile_hdr &&
                     b->blk_tile_count && b->blk_tiles,
                 "MPM batch: NULL buffer");
    return 0;
}

int check_horizon_save(const dl_mpm_batch *b) {
    const int E = b->params.dim == 3 ? 4 : 8;
    int nblk = 1;
    for (int d = 0; d < b->params.dim; ++d) nblk *= b->params.n_grid / E;
    DL_CHECK_ARG(b->hgrid && b->hvo && b->hlist && b->hcount && b->hover && b->hcap > 0 &&
                     b->hcap <= nblk,
                 "MPM batch: the horizon grid save needs hgrid, hvo, hlist, hcount, hover and "
                 "0 < hcap <= %d", nblk);
    return 0;
}

int check_control(const dl_mpm_batch *b) {
    DL_CHECK_ARG(b->params.dim == 3, "MPM controller: 3D robots only");
    DL_CHECK_ARG(b->n_groups > 0 && b->n_sin >= 0 && b->run_period > 0 && b->n_hidden > 0 &&
                     b->offset_scale != 0 && b->max_speed != 0,
                 "MPM controller: bad layout");
    DL_CHECK_ARG(b->group && b->inv_count && b->targets && b->W1 && b->b1 && b->W2 && b->b2 &&
                     b->sens && b->center && b->vcom && b->inp && b->hid,
                 "MPM controller: NULL buffer");
    return 0;
}

int n_in(const dl_mpm_batch *b) { return b->n_sin + 3 + 3 * b->n_groups + 4; }

template <typename T> Ctrl<T> make_ctrl(const dl_mpm_batch *b) {
    Ctrl<T> c;
    c.P = b->P; c.Gs = b->n_groups; c.n_sin = b->n_sin; c.run_period = b->run_period;
    c.I = n_in(b); c.H = b->n_hidden; c.O = b->n_act; c.act = b->act;
    c.scale = T(b->offset_scale); c.max_speed = T(b->max_speed);
    c.inv_count = (const T *)b->inv_count; c.targets = (const T *)b->targets;
    c.W1 = (const T *)b->W1; c.b1 = (const T *)b->b1; c.W2 = (const T *)b->W2;
    c.b2 = (const T *)b->b2;
    c.np = b->n_part;
    c.hz = b->horizons;
    c.per_robot_weights = b->per_robot_weights;
    c.t = 0;
    return c;
}

template <int D> Blk make_blk(const dl_mpm_batch *b) {
    Blk k;
    k.flag = b->blk_flag; k.list = b->blk_list; k.count = b->blk_count;
    k.nb = b->params.n_grid / BlkE<D>::E;
    k.nblk = 1;
    for (int d = 0; d < D; ++d) k.nblk *= k.nb;
    return k;
}

template <int D, typename T> struct Slot {
    T *x, *v, *C, *F;
};
template <int D, typename T> Slot<D, T> slot(const dl_mpm_slots *s, int i, int B, int P) {
    const size_t sv = (size_t)B * D * P, sm = (size_t)B * D * D * P;
    return {(T *)s->x + i * sv, (T *)s->v + i * sv, (T *)s->C + i * sm, (T *)s->F + i * sm};
}

template <int D, typename T> GridSave<T> save_at(const dl_mpm_batch *b, int j) {
    if (!b->grid_save) return {nullptr, nullptr, nullptr, nullptr, 0, nullptr};
    const Blk bk = make_blk<D>(b);
    const size_t rec = (size_t)b->B * bk.nblk;
    return {(T *)b->grid_save + j * rec * 64 * 4, (T *)b->vo_save + j * rec * 64 * 4,
            b->list_save + j * rec, b->count_save + (size_t)j * b->B, bk.nblk, nullptr};
}

// Horizon save of step t (hcap blocks per robot), overflow flag of step t.
template <typename T> GridSave<T> hsave_at(const dl_mpm_batch *b, int t) {
    const size_t rec = (size_t)b->B * b->hcap;
    return {(T *)b->hgrid + t * rec * 64 * 4, (T *)b->hvo + t * rec * 64 * 4,
            b->hlist + t * rec, b->hcount + (size_t)t * b->B, b->hcap, b->hover + t};
}

int n_windows(const dl_mpm_batch *b) {
    return b->windows ? b->n_windows : (b->P + PB - 1) / PB;
}

template <typename T> TileBuf<T> make_tiles(const dl_mpm_batch *b) {
    return {(T *)b->tiles, b->tile_hdr, b->blk_tile_count, b->blk_tiles, n_windows(b),
            b->tile_nodes, b->scatter == DL_MPM_SCATTER_STORE};
}

int grid_wgs(const Blk &k) { return std::max(1, std::min(64, (k.nblk + NWAVE - 1) / NWAVE)); }

template <int D, typename T>
int forward_impl(const dl_mpm_batch *b, const dl_mpm_slots *st, int t0, int n, int flags,
                 hipStream_t s) {
    const int B = b->B, P = b->P, O = b->n_act, NS = 3 * b->n_groups + 6;
    const bool control = flags & DL_MPM_CONTROL;
    MP<T> mp = make_mp<T>(b->params);
    mp.np = b->n_part;
    mp.hz = b->horizons;
    const Blk bk = make_blk<D>(b);
    const dim3 pg((P + PB - 1) / PB, B), gg(grid_wgs(bk), B), pw(n_windows(b), B);
    const int I = n_in(b), H = b->n_hidden;
    const T *act = (const T *)b->act_out;
    const TileBuf<T> tb = make_tiles<T>(b);
    const size_t tile_bytes = (size_t)tb.tn * 4 * sizeof(T);
    const int unroll = b->scatter_unroll;
    DL_UNROLL_DISPATCH(U_, DL_CHECK_HIP(ensure_dynamic_lds((const void *)k_p2g_b<D, T, true, U_>,
                                                           tile_bytes)));
    const Ctrl<T> ctrl = make_ctrl<T>(b);
    if constexpr (D == 3) {
        if (control) {
            DL_CHECK_ARG((size_t)(I + H + O) * sizeof(T) <= tile_bytes,
                         "MPM batch: the controller's LDS rows exceed the scatter tile");
            DL_UNROLL_DISPATCH(U_, DL_CHECK_HIP(ensure_dynamic_lds(
                                       (const void *)k_p2g_b<D, T, true, U_, true>, tile_bytes)));
        }
    }
    if (control && (flags & DL_MPM_SENSE_FIRST)) {
        const Slot<D, T> s0 = slot<D, T>(st, 0, B, P);
        if constexpr (D == 3)
            hipLaunchKernelGGL(k_sense<T>, pg, dim3(PB), NS * sizeof(T), s, P, b->n_part, s0.x, s0.v,
                               SenseOut<T>{(T *)b->sens + (size_t)t0 * B * NS, b->group,
                                           b->n_groups});
    }
    for (int j = 0; j < n; ++j) {
        const int t = t0 + j;
        mp.t = t;
        const Slot<D, T> a = slot<D, T>(st, j % st->n_slots, B, P);
        const Slot<D, T> c = slot<D, T>(st, (j + 1) % st->n_slots, B, P);
        if (flags & DL_MPM_REPLAY) {
            hipLaunchKernelGGL((k_grid_restore<D, T>), gg, dim3(PB), 0, s, mp, (T *)b->grid,
                               (T *)b->grid_vo, bk, hsave_at<T>(b, t), 1);
            hipLaunchKernelGGL((k_g2p_b<D, T, false, true>), pg, dim3(PB), 0, s, mp, P, a.x,
                               (const T *)b->grid_vo, c.x, c.v, c.C, b->blk_count,
                               SenseOut<T>{nullptr, nullptr, 0}, (const T *)a.C,
                               (const T *)a.F, c.F);
            continue;
        }
        PView<D, T> pv{a.x, a.v, a.C, a.F, P};
        bool fused = false;
        if constexpr (D == 3) {
            if (control) {
                // the step's controller runs inside P2G (no launch of its own)
                const CtrlRows<T> cr{(const T *)b->sens + (size_t)t * B * NS,
                                     (T *)b->center + (size_t)t * B * 3,
                                     (T *)b->vcom + (size_t)t * B * 3,
                                     (T *)b->inp + (size_t)t * B * I,
                                     (T *)b->hid + (size_t)t * B * H,
                                     (T *)b->act_out + (size_t)t * B * O};
                DL_UNROLL_DISPATCH(U_, hipLaunchKernelGGL((k_p2g_b<D, T, true, U_, true>), pw,
                                                          dim3(PB), tile_bytes, s, mp, pv,
                                                          act + (size_t)t * B * O, b->act_id, O,
                                                          c.F, (T *)b->grid, bk, b->err, tb,
                                                          b->windows, ctrl, cr));
                fused = true;
            }
        }
        if (!fused)
            DL_UNROLL_DISPATCH(U_, hipLaunchKernelGGL((k_p2g_b<D, T, true, U_>), pw, dim3(PB),
                                                      tile_bytes, s, mp, pv, act + (size_t)t * B * O,
                                                      b->act_id, O, c.F, (T *)b->grid, bk, b->err,
                                                      tb, b->windows));
        const GridSave<T> sv = (flags & DL_MPM_SAVE_GRID)      ? save_at<D, T>(b, j)
                               : (flags & DL_MPM_SAVE_HORIZON) ? hsave_at<T>(b, t)
                                                               : GridSave<T>{nullptr, nullptr, nullptr, nullptr, 0, nullptr};
        hipLaunchKernelGGL((k_grid_b<D, T>), gg, dim3(PB), 0, s, mp, (T *)b->grid, (T *)b->grid_vo,
                           bk, 1, tb, sv);
        if (control && D == 3)
            hipLaunchKernelGGL((k_g2p_b<D, T, true>), pg, dim3(PB), NS * sizeof(T), s, mp, P, a.x,
                               (const T *)b->grid_vo, c.x, c.v, c.C, b->blk_count,
                               SenseOut<T>{(T *)b->sens + (size_t)(t + 1) * B * NS, b->group,
                                           b->n_groups});
        else
            hipLaunchKernelGGL((k_g2p_b<D, T, false>), pg, dim3(PB), 0, s, mp, P, a.x,
                               (const T *)b->grid_vo, c.x, c.v, c.C, b->blk_count,
                               SenseOut<T>{nullptr, nullptr, 0});
    }
    if (control && t0 + n == b->horizon)
        hipLaunchKernelGGL(k_center<T>, dim3(B), dim3(64), 0, s, ctrl, t0 + n,
                           (const T *)b->sens + (size_t)(t0 + n) * B * NS,
                           (T *)b->center + (size_t)(t0 + n) * B * 3,
                           (T *)b->vcom + (size_t)(t0 + n) * B * 3);
    DL_CHECK_LAUNCH();
    return 0;
}

template <int D, typename T>
int backward_impl(const dl_mpm_batch *b, const dl_mpm_slots *seg, const dl_mpm_slots *adj, int t0,
                  int n, hipStream_t s) {
    const int B = b->B, P = b->P, O = b->n_act;
    MP<T> mp = make_mp<T>(b->params);
    mp.np = b->n_part;
    mp.hz = b->horizons;
    const Blk bk = make_blk<D>(b);
    const dim3 pg((P + PB - 1) / PB, B), gg(grid_wgs(bk), B), pw(n_windows(b), B);
    const int I = n_in(b), H = b->n_hidden;
    Ctrl<T> ctrl = make_ctrl<T>(b);
    const T *act = (const T *)b->act_out;
    const TileBuf<T> tb = make_tiles<T>(b);
    const size_t tile_bytes = (size_t)tb.tn * 4 * sizeof(T);
    const int unroll = b->scatter_unroll;
    DL_UNROLL_DISPATCH(U_, DL_CHECK_HIP(ensure_dynamic_lds((const void *)k_p2g_b<D, T, false, U_>,
                                                           tile_bytes)));
    DL_UNROLL_DISPATCH(U_, DL_CHECK_HIP(ensure_dynamic_lds((const void *)k_g2p_adj_b<D, T, U_>,
                                                           tile_bytes)));
    const bool restore = b->restore_from == 1 || b->grid_save;
    auto saved = [&](int jj) {
        return b->restore_from == 1 ? hsave_at<T>(b, t0 + jj) : save_at<D, T>(b, jj);
    };
    int cur = 0;
    bool restored = false;  // step j's grid restored by the previous launch
    for (int j = n - 1; j >= 0; --j) {
        const int t = t0 + j, nxt = 1 - cur;
        mp.t = t;
        ctrl.t = t;
        const Slot<D, T> a = slot<D, T>(seg, j, B, P);
        const Slot<D, T> A1 = slot<D, T>(adj, cur, B, P), A0 = slot<D, T>(adj, nxt, B, P);
        PView<D, T> pv{a.x, a.v, a.C, a.F, P};
        if (restore) {
            if (!restored)
                hipLaunchKernelGGL((k_grid_restore<D, T>), gg, dim3(PB), 0, s, mp, (T *)b->grid,
                                   (T *)b->grid_vo, bk, saved(j), 0);
        } else {
            DL_UNROLL_DISPATCH(U_, hipLaunchKernelGGL((k_p2g_b<D, T, false, U_>), pw, dim3(PB),
                                                      tile_bytes, s, mp, pv,
                                                      act + (size_t)t * B * O, b->act_id, O,
                                                      (T *)nullptr, (T *)b->grid, bk, b->err, tb,
                                                      b->windows));
            hipLaunchKernelGGL((k_grid_b<D, T>), gg, dim3(PB), 0, s, mp, (T *)b->grid,
                               (T *)b->grid_vo, bk, 0, tb,
                               GridSave<T>{nullptr, nullptr, nullptr, nullptr, 0, nullptr});
        }
        SenseAdj<T> sa{};
        if (D == 3 && b->in_grad) {
            sa.in_b = (const T *)b->in_grad + (size_t)(t + 1) * B * I;
            sa.cg = (const T *)b->center_grad + (size_t)(t + 1) * B * 3;
            sa.vg = (const T *)b->vcom_grad + (size_t)(t + 1) * B * 3;
            sa.group = b->group;
            sa.inv_count = (const T *)b->inv_count;
            sa.Gs = b->n_groups;
            sa.n_sin = b->n_sin;
            sa.I = I;
            sa.scale = T(b->offset_scale);
        }
        DL_UNROLL_DISPATCH(U_, hipLaunchKernelGGL((k_g2p_adj_b<D, T, U_>), pw, dim3(PB),
                                                  tile_bytes, s, mp, P, a.x,
                                                  (const T *)b->grid_vo, A1.x, A1.v, A1.C, A0.x,
                                                  (T *)b->grid_vo_grad, sa, bk, tb, b->windows));
        hipLaunchKernelGGL((k_grid_adj_b<D, T>), gg, dim3(PB), 0, s, mp, (T *)b->grid,
                           (T *)b->grid_vo_grad, (T *)b->grid_grad, bk, tb);
        hipLaunchKernelGGL((k_p2g_adj_b<D, T>), pg, dim3(PB), O * sizeof(T), s, mp, pv,
                           act + (size_t)t * B * O, b->act_id, O, (const T *)b->grid_grad, A1.F,
                           A0.x, A0.v, A0.C, A0.F, (T *)b->act_grad, b->blk_count);
        restored = false;
        if (b->in_grad && restore && j > 0) {
            hipLaunchKernelGGL((k_restore_ctrl_adj<D, T>), dim3(gg.x + 1, B), dim3(PB),
                               (I + 3 * H + O) * sizeof(T), s, mp, (T *)b->grid, (T *)b->grid_vo,
                               bk, saved(j - 1), (int)gg.x, ctrl,
                               (const T *)b->inp + (size_t)t * B * I, (T *)b->act_grad,
                               (T *)b->in_grad + (size_t)t * B * I,
                               (T *)b->hid_grad + (size_t)t * B * H,
                               (T *)b->out_grad + (size_t)t * B * O);
            restored = true;
        } else if (b->in_grad) {
            hipLaunchKernelGGL(k_ctrl_adj<T>, dim3(B), dim3(64), (I + 3 * H + O) * sizeof(T), s,
                               ctrl, (const T *)b->inp + (size_t)t * B * I, (T *)b->act_grad,
                               (T *)b->in_grad + (size_t)t * B * I,
                               (T *)b->hid_grad + (size_t)t * B * H,
                               (T *)b->out_grad + (size_t)t * B * O);
        }
        cur = nxt;
    }
    if (b->in_grad) {
        const int nw = H * I + H + O * H + O;
        hipLaunchKernelGGL(k_wgrad_time<T>, dim3((nw + 255) / 256, B), dim3(256), 0, s, B, I, H,
                           O, n, (const T *)b->inp + (size_t)t0 * B * I,
                           (const T *)b->hid + (size_t)t0 * B * H,
                           (const T *)b->hid_grad + (size_t)t0 * B * H,
                           (const T *)b->out_grad + (size_t)t0 * B * O, (T *)b->wgrad,
                           b->horizons, t0);
    }
    if (cur != 0) {
        const Slot<D, T> from = slot<D, T>(adj, 1, B, P), to = slot<D, T>(adj, 0, B, P);
        const size_t nv = (size_t)B * D * P, nm = (size_t)B * D * D * P;
        hipLaunchKernelGGL(k_copy<T>, dim3(1024), dim3(256), 0, s, from.x, to.x, nv);
        hipLaunchKernelGGL(k_copy<T>, dim3(1024), dim3(256), 0, s, from.v, to.v, nv);
        hipLaunchKernelGGL(k_copy<T>, dim3(1024), dim3(256), 0, s, from.C, to.C, nm);
        hipLaunchKernelGGL(k_copy<T>, dim3(1024), dim3(256), 0, s, from.F, to.F, nm);
    }
    DL_CHECK_LAUNCH();
    return 0;
}

}  // namespace

#define MPMB_DISPATCH(b, FN, ...)                                                          \
    do {                                                                                   \
        if ((b)->params.dim == 3 && (b)->dtype == DL_F32) return FN<3, float>(__VA_ARGS__);  \
        if ((b)->params.dim == 3 && (b)->dtype == DL_F64) return FN<3, double>(__VA_ARGS__); \
        if ((b)->params.dim == 2 && (b)->dtype == DL_F32) return FN<2, float>(__VA_ARGS__);  \
        return FN<2, double>(__VA_ARGS__);                                                 \
    } while (0)

extern "C" {

int32_t dl_mpm_forward(const dl_mpm_batch *b, const dl_mpm_slots *st, int32_t t0, int32_t n,
                       int32_t flags, void *stream) {
    if (int e = check_batch(b)) return e;
    DL_CHECK_ARG(st && st->x && st->v && st->C && st->F && st->n_slots >= 2,
                 "dl_mpm_forward: bad state slots");
    DL_CHECK_ARG(n > 0 && t0 >= 0 && t0 + n <= b->horizon, "dl_mpm_forward: steps [%d, %d) "
                 "outside the horizon %d", t0, t0 + n, b->horizon);
    if (flags & DL_MPM_CONTROL)
        if (int e = check_control(b)) return e;
    if (flags & DL_MPM_SAVE_GRID)
        DL_CHECK_ARG(b->grid_save && b->vo_save && b->list_save && b->count_save &&
                         n <= b->save_steps,
                     "dl_mpm_forward: DL_MPM_SAVE_GRID needs the save buffers for %d steps", n);
    if (flags & (DL_MPM_SAVE_HORIZON | DL_MPM_REPLAY))
        if (int e = check_horizon_save(b)) return e;
    DL_CHECK_ARG(!((flags & DL_MPM_SAVE_GRID) && (flags & DL_MPM_SAVE_HORIZON)),
                 "dl_mpm_forward: DL_MPM_SAVE_GRID and DL_MPM_SAVE_HORIZON exclude each other");
    DL_CHECK_ARG(!(flags & DL_MPM_REPLAY) ||
                     !(flags & (DL_MPM_CONTROL | DL_MPM_SENSE_FIRST | DL_MPM_SAVE_GRID |
                                DL_MPM_SAVE_HORIZON)),
                 "dl_mpm_forward: DL_MPM_REPLAY recomputes only (no controller, no saving)");
    MPMB_DISPATCH(b, forward_impl, b, st, t0, n, flags, (hipStream_t)stream);
}

int32_t dl_mpm_backward(const dl_mpm_batch *b, const dl_mpm_slots *seg, const dl_mpm_slots *adj,
                        int32_t t0, int32_t n, void *stream) {
    if (int e = check_batch(b)) return e;
    DL_CHECK_ARG(seg && seg->x && seg->v && seg->C && seg->F && seg->n_slots >= n + 1,
                 "dl_mpm_backward: the segment needs n + 1 state slots");
    DL_CHECK_ARG(adj && adj->x && adj->v && adj->C && adj->F && adj->n_slots == 2,
                 "dl_mpm_backward: the adjoint needs 2 slots");
    DL_CHECK_ARG(n > 0 && t0 >= 0 && t0 + n <= b->horizon, "dl_mpm_backward: bad step range");
    DL_CHECK_ARG(b->grid_vo_grad && b->grid_grad && b->act_grad, "dl_mpm_backward: NULL buffer");
    DL_CHECK_ARG(!b->grid_save || (b->vo_save && b->list_save && b->count_save &&
                                   n <= b->save_steps),
                 "dl_mpm_backward: saved grids cover %d steps, segment has %d", b->save_steps, n);
    DL_CHECK_ARG(b->restore_from == 0 || b->restore_from == 1,
                 "dl_mpm_backward: restore_from must be 0 or 1");
    if (b->restore_from == 1)
        if (int e = check_horizon_save(b)) return e;
    if (b->in_grad) {
        if (int e = check_control(b)) return e;
        DL_CHECK_ARG(b->center_grad && b->vcom_grad && b->wgrad && b->hid_grad && b->out_grad,
                     "dl_mpm_backward: NULL buffer");
    }
    MPMB_DISPATCH(b, backward_impl, b, seg, adj, t0, n, (hipStream_t)stream);
}

}  // extern "C"
