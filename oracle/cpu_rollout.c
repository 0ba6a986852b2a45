/* CPU restatement of the quadruped training rollout in C with OpenMP:
 * forward, hand-written adjoint, checkpoints every K steps with each segment
 * recomputed in the backward, one Adam step.
 *
 * TEST INFRASTRUCTURE AND CPU BASELINE ONLY.  Nothing in the product package
 * loads this library; only tests/ (compared against the torch oracle)
 * and the cpu_baseline leg of bench.py (as the timed CPU figure) do.
 *
 * Built twice by oracle/build_cpu.py: REAL = float (the baseline) and
 * REAL = double (pinned against the FP64 torch oracle, whose autograd stands
 * in for the reference's generated adjoints).
 *
 * Reference functions restated (file:line in the reference repo):
 *   springs()     SolverMassSpring.apply_spring_force  multitask/solver_mass_spring.py:86-105
 *   toi()         SolverMassSpring.advance_toi         multitask/solver_mass_spring.py:117-148
 *   center        SolverMassSpring.compute_center      multitask/solver_mass_spring.py:150-156
 *   sensors()     TaichiEnv.nn_input                   multitask/taichi_env.py:149-183
 *                 (the quadruped workload's sensor layout: 10 sinusoids, COM
 *                 velocity, vertex offsets / 0.2, target velocity / 0.08, target height)
 *   mlp           Model.nn1 / nn2                      multitask/nn.py:130-150
 *   the adjoint   ti.ad.Tape over the above            multitask/diffphy_trainer.py:245-256
 *   cr_adam()     Model.adam_update                    multitask/nn.py:184-191
 * The loss is the quadruped workload's: mean over the last W states of
 * |v_com,xz - v_target,xz|^2 + (c_y - h_target)^2, averaged over tasks
 * (oracle/mass_spring.py rollout_task).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#ifndef REAL
#define REAL float
#endif
typedef REAL real;

/* math in the precision of real */
#define R_FN(name, fn) \
    static inline real name(real x) { return _Generic(x, float: fn##f, default: fn)(x); }
R_FN(r_sqrt, sqrt)
R_FN(r_tanh, tanh)
R_FN(r_sin, sin)
R_FN(r_cos, cos)
R_FN(r_exp, exp)
static inline real r_pow(real x, real y) { return _Generic(x, float: powf, default: pow)(x, y); }

typedef struct {
    int32_t N, S, O, H, I;
    const int32_t *a, *b, *act_index;  /* [S] */
    const real *L0, *k, *act_coeff;    /* [S] */
    real dt, gravity, ground, drag, dashpot, friction, stiffness_scale;
    int32_t n_sin, run_period;
    real offset_scale, max_speed;
    int32_t act_tanh;                  /* 1: tanh, 0: sin */
    int32_t T, W, K;                   /* horizon, loss window, checkpoint interval */
} cr_cfg;

static inline real actf(int tanh_, real z) { return tanh_ ? r_tanh(z) : r_sin(z); }
static inline real dactf(int tanh_, real z, real y) { return tanh_ ? (real)1 - y * y : r_cos(z); }

/* per-task working memory */
typedef struct {
    real *ck;                     /* checkpoints [nseg][6N] */
    real *xs, *vs;                /* segment states [K][3N] */
    real *inp, *z1, *h, *z2, *y;  /* segment caches [K][I], [K][H], [K][H], [K][O], [K][O] */
    real *vincs;                  /* segment impulse sums [K][3N] */
    real *x, *v, *xn, *vn;        /* [3N] */
    real *xb, *vb, *xb1, *vb1, *vincb, *oldvb;
    real *inpb, *z1b, *hb, *z2b, *ab;
} work_t;

static int work_alloc(work_t *w, const cr_cfg *c) {
    const size_t N3 = 3 * (size_t)c->N, K = c->K, nseg = (c->T + c->K - 1) / c->K;
    memset(w, 0, sizeof(*w));
    w->ck = malloc(sizeof(real) * nseg * 2 * N3);
    w->xs = malloc(sizeof(real) * K * N3);
    w->vs = malloc(sizeof(real) * K * N3);
    w->inp = malloc(sizeof(real) * K * c->I);
    w->z1 = malloc(sizeof(real) * K * c->H);
    w->h = malloc(sizeof(real) * K * c->H);
    w->z2 = malloc(sizeof(real) * K * c->O);
    w->y = malloc(sizeof(real) * K * c->O);
    w->vincs = malloc(sizeof(real) * K * N3);
    real **v3[] = {&w->x, &w->v, &w->xn, &w->vn, &w->xb, &w->vb, &w->xb1, &w->vb1,
                   &w->vincb, &w->oldvb};
    for (size_t i = 0; i < sizeof(v3) / sizeof(v3[0]); ++i) *v3[i] = malloc(sizeof(real) * N3);
    w->inpb = malloc(sizeof(real) * c->I);
    w->z1b = malloc(sizeof(real) * c->H);
    w->hb = malloc(sizeof(real) * c->H);
    w->z2b = malloc(sizeof(real) * c->O);
    w->ab = malloc(sizeof(real) * c->O);
    return w->ck && w->xs && w->vs && w->inp && w->z1 && w->h && w->z2 && w->y && w->x && w->v &&
           w->xn && w->vn && w->vincs && w->xb && w->vb && w->xb1 && w->vb1 && w->vincb &&
           w->oldvb && w->inpb && w->z1b && w->hb && w->z2b && w->ab;
}

static void work_free(work_t *w) {
    void *p[] = {w->ck, w->xs, w->vs, w->inp, w->z1, w->h, w->z2, w->y, w->x, w->v, w->xn,
                 w->vn, w->vincs, w->xb, w->vb, w->xb1, w->vb1, w->vincb, w->oldvb, w->inpb,
                 w->z1b, w->hb, w->z2b, w->ab};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); ++i) free(p[i]);
}

static void means(const cr_cfg *c, const real *x, const real *v, real cm[3], real vc[3]) {
    real sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    const real inv = (real)1 / (real)c->N;
    for (int i = 0; i < c->N; ++i)
        for (int d = 0; d < 3; ++d) {
            sx[d] += x[3 * i + d] * inv;
            sv[d] += v[3 * i + d] * inv;
        }
    for (int d = 0; d < 3; ++d) {
        cm[d] = sx[d];
        vc[d] = sv[d];
    }
}

static void sensors(const cr_cfg *c, int t, const real *x, const real cm[3], const real vc[3],
                    const real tv[3], real th, real *inp) {
    const double pi2 = 6.283185307179586;
    int q = 0;
    for (int j = 0; j < c->n_sin; ++j)
        inp[q++] = (real)sin(pi2 / c->run_period * t + pi2 / c->n_sin * j);
    for (int d = 0; d < 3; ++d) inp[q++] = vc[d];
    for (int i = 0; i < c->N; ++i)
        for (int d = 0; d < 3; ++d) inp[q++] = (x[3 * i + d] - cm[d]) / c->offset_scale;
    for (int d = 0; d < 3; ++d) inp[q++] = tv[d] / c->max_speed;
    inp[q++] = th;
}

static void mlp(const cr_cfg *c, const real *W1, const real *b1, const real *W2, const real *b2,
                const real *inp, real *z1, real *h, real *z2, real *y) {
    for (int j = 0; j < c->H; ++j) {
        const real *wr = W1 + (size_t)j * c->I;
        real acc = 0;
#pragma omp simd reduction(+ : acc)
        for (int i = 0; i < c->I; ++i) acc += wr[i] * inp[i];
        z1[j] = acc + b1[j];
        h[j] = actf(c->act_tanh, z1[j]);
    }
    for (int o = 0; o < c->O; ++o) {
        const real *wr = W2 + (size_t)o * c->H;
        real acc = 0;
#pragma omp simd reduction(+ : acc)
        for (int j = 0; j < c->H; ++j) acc += wr[j] * h[j];
        z2[o] = acc + b2[o];
        y[o] = actf(c->act_tanh, z2[o]);
    }
}

/* v_inc of the spring and dashpot impulses */
static void springs(const cr_cfg *c, const real *x, const real *v, const real *act, real *vinc) {
    memset(vinc, 0, sizeof(real) * 3 * c->N);
    const real ks = c->stiffness_scale;
    for (int s = 0; s < c->S; ++s) {
        const int ia = c->a[s], ib = c->b[s];
        real dist[3], nsq = 0;
        for (int d = 0; d < 3; ++d) {
            dist[d] = x[3 * ia + d] - x[3 * ib + d];
            nsq += dist[d] * dist[d];
        }
        const real len = r_sqrt(nsq + (real)1e-8) + (real)1e-4;
        const real target = c->L0[s] * ((real)1 + c->act_coeff[s] * act[c->act_index[s]]);
        const real f = c->dt * (len - target) * (c->k[s] * ks) / len;
        const real rn = (real)1 / r_sqrt(nsq);
        real vrel = 0;
        for (int d = 0; d < 3; ++d) vrel += (v[3 * ia + d] - v[3 * ib + d]) * dist[d] * rn;
        for (int d = 0; d < 3; ++d) {
            const real imp = f * dist[d] + c->dashpot * vrel * dist[d] * rn;
            vinc[3 * ia + d] -= imp;
            vinc[3 * ib + d] += imp;
        }
    }
}

static inline real fric(real cv, real fr) {
    return cv < 0 ? (cv + fr < 0 ? cv + fr : 0) : (cv - fr > 0 ? cv - fr : 0);
}

/* semi-implicit Euler with time-of-impact ground contact and friction */
static void toi(const cr_cfg *c, const real *x, const real *v, const real *vinc, real *xn,
                real *vn) {
    const real sdrag = r_exp(-c->dt * c->drag), g = c->dt * c->gravity;
    for (int i = 0; i < c->N; ++i) {
        real ov[3];
        for (int d = 0; d < 3; ++d) ov[d] = sdrag * v[3 * i + d] + vinc[3 * i + d];
        ov[1] += g;
        const real ny = x[3 * i + 1] + c->dt * ov[1];
        const int hit = ny < c->ground && ov[1] < (real)-1e-4;
        if (!hit) {
            for (int d = 0; d < 3; ++d) {
                vn[3 * i + d] = ov[d];
                xn[3 * i + d] = x[3 * i + d] + c->dt * ov[d];
            }
            continue;
        }
        const real t = -(x[3 * i + 1] - c->ground) / ov[1], fr = c->friction * -ov[1];
        real nv[3] = {fric(ov[0], fr), 0, fric(ov[2], fr)};
        for (int d = 0; d < 3; ++d) {
            vn[3 * i + d] = nv[d];
            xn[3 * i + d] = x[3 * i + d] + t * ov[d] + (c->dt - t) * nv[d];
        }
    }
}

/* one step forward from (x, v) at time t: keeps the MLP values and the
 * impulse sums in slot j */
static void step_fwd(const cr_cfg *c, work_t *w, int j, int t, const real *x, const real *v,
                     const real tv[3], real th, const real *W1, const real *b1, const real *W2,
                     const real *b2, real *xn, real *vn) {
    real cm[3], vc[3];
    means(c, x, v, cm, vc);
    real *inp = w->inp + (size_t)j * c->I, *z1 = w->z1 + (size_t)j * c->H;
    real *h = w->h + (size_t)j * c->H, *z2 = w->z2 + (size_t)j * c->O, *y = w->y + (size_t)j * c->O;
    sensors(c, t, x, cm, vc, tv, th, inp);
    mlp(c, W1, b1, W2, b2, inp, z1, h, z2, y);
    real *vinc = w->vincs + (size_t)j * 3 * c->N;
    springs(c, x, v, y, vinc);
    toi(c, x, v, vinc, xn, vn);
}

static real loss_term(const real cm[3], const real vc[3], const real tv[3], real th) {
    const real a = vc[0] - tv[0], b = vc[2] - tv[2], e = cm[1] - th;
    return a * a + b * b + e * e;
}

/* adds the loss term's adjoint at a state (scale = dL/d term) to xb, vb */
static void loss_adj(const cr_cfg *c, const real *x, const real *v, const real tv[3], real th,
                     real scale, real *xb, real *vb) {
    real cm[3], vc[3];
    means(c, x, v, cm, vc);
    const real inv = (real)1 / (real)c->N;
    const real gvx = 2 * (vc[0] - tv[0]) * scale * inv, gvz = 2 * (vc[2] - tv[2]) * scale * inv;
    const real gcy = 2 * (cm[1] - th) * scale * inv;
    for (int i = 0; i < c->N; ++i) {
        vb[3 * i + 0] += gvx;
        vb[3 * i + 2] += gvz;
        xb[3 * i + 1] += gcy;
    }
}

/* adjoint of step t (state t -> t+1): given xb1, vb1 = dL/d(state t+1),
 * writes xb, vb = the part of dL/d(state t) through this step, and adds the
 * weight gradients into g */
static void step_adj(const cr_cfg *c, work_t *w, int j, int t, const real *x, const real *v,
                     const real *W1, const real *W2, real *g) {
    const int N = c->N, H = c->H, I = c->I, O = c->O;
    const real *inp = w->inp + (size_t)j * I, *z1 = w->z1 + (size_t)j * H;
    const real *h = w->h + (size_t)j * H, *z2 = w->z2 + (size_t)j * O, *y = w->y + (size_t)j * O;
    (void)t;
    const real *vinc = w->vincs + (size_t)j * 3 * N;   /* kept by the recompute */
    /* toi adjoint */
    const real sdrag = r_exp(-c->dt * c->drag), gv = c->dt * c->gravity;
    real *xb = w->xb, *vb = w->vb;
    for (int i = 0; i < N; ++i) {
        real ov[3];
        for (int d = 0; d < 3; ++d) ov[d] = sdrag * v[3 * i + d] + vinc[3 * i + d];
        ov[1] += gv;
        const real ny = x[3 * i + 1] + c->dt * ov[1];
        const int hit = ny < c->ground && ov[1] < (real)-1e-4;
        const real *xb1 = w->xb1 + 3 * i, *vb1 = w->vb1 + 3 * i;
        real ob[3];
        for (int d = 0; d < 3; ++d) xb[3 * i + d] = xb1[d];
        if (!hit) {
            for (int d = 0; d < 3; ++d) ob[d] = vb1[d] + c->dt * xb1[d];
        } else {
            const real tt = -(x[3 * i + 1] - c->ground) / ov[1], fr = c->friction * -ov[1];
            const real nv[3] = {fric(ov[0], fr), 0, fric(ov[2], fr)};
            real tb = 0, frb = 0;
            for (int d = 0; d < 3; ++d) {
                tb += xb1[d] * (ov[d] - nv[d]);
                ob[d] = tt * xb1[d];
            }
            for (int d = 0; d < 3; d += 2) {
                const real nvb = vb1[d] + (c->dt - tt) * xb1[d];
                const real cv = ov[d];
                if (cv < 0) {
                    if (cv + fr < 0) {
                        ob[d] += nvb;
                        frb += nvb;
                    }
                } else if (cv - fr > 0) {
                    ob[d] += nvb;
                    frb -= nvb;
                }
            }
            ob[1] += -c->friction * frb;
            xb[3 * i + 1] += -tb / ov[1];
            ob[1] += tb * (x[3 * i + 1] - c->ground) / (ov[1] * ov[1]);
        }
        for (int d = 0; d < 3; ++d) {
            vb[3 * i + d] = sdrag * ob[d];
            w->vincb[3 * i + d] = ob[d];
        }
    }
    /* spring adjoint */
    real *ab = w->ab;
    for (int o = 0; o < O; ++o) ab[o] = 0;
    const real ks = c->stiffness_scale;
    for (int s = 0; s < c->S; ++s) {
        const int ia = c->a[s], ib = c->b[s];
        real dist[3], dv[3], nsq = 0;
        for (int d = 0; d < 3; ++d) {
            dist[d] = x[3 * ia + d] - x[3 * ib + d];
            dv[d] = v[3 * ia + d] - v[3 * ib + d];
            nsq += dist[d] * dist[d];
        }
        const real sq = r_sqrt(nsq + (real)1e-8), len = sq + (real)1e-4;
        const real target = c->L0[s] * ((real)1 + c->act_coeff[s] * y[c->act_index[s]]);
        const real kk = c->k[s] * ks, f = c->dt * (len - target) * kk / len;
        const real rn = (real)1 / r_sqrt(nsq);
        real dn[3], vrel = 0, ib_[3];
        for (int d = 0; d < 3; ++d) {
            dn[d] = dist[d] * rn;
            vrel += dv[d] * dn[d];
            ib_[d] = w->vincb[3 * ib + d] - w->vincb[3 * ia + d];   /* d L / d imp */
        }
        /* imp = f dist + dashpot vrel dn */
        real vrelb = 0, fb = 0, distb[3], dnb[3];
        for (int d = 0; d < 3; ++d) {
            vrelb += c->dashpot * ib_[d] * dn[d];
            fb += ib_[d] * dist[d];
        }
        for (int d = 0; d < 3; ++d) {
            dnb[d] = c->dashpot * vrel * ib_[d] + vrelb * dv[d];
            distb[d] = f * ib_[d] + rn * dnb[d];
            vb[3 * ia + d] += vrelb * dn[d];
            vb[3 * ib + d] -= vrelb * dn[d];
        }
        real rnb = 0;
        for (int d = 0; d < 3; ++d) rnb += dnb[d] * dist[d];
        const real lenb = fb * c->dt * kk * target / (len * len);
        const real targetb = -fb * c->dt * kk / len;
        const real nsqb = lenb * (real)0.5 / sq - (real)0.5 * rnb * rn * rn * rn;
        for (int d = 0; d < 3; ++d) {
            const real gd = distb[d] + 2 * nsqb * dist[d];
            xb[3 * ia + d] += gd;
            xb[3 * ib + d] -= gd;
        }
        ab[c->act_index[s]] += targetb * c->L0[s] * c->act_coeff[s];
    }
    /* MLP adjoint; g = [W1 (H I) | b1 (H) | W2 (O H) | b2 (O)] */
    real *gW1 = g, *gb1 = g + (size_t)H * I, *gW2 = gb1 + H, *gb2 = gW2 + (size_t)O * H;
    for (int o = 0; o < O; ++o) {
        w->z2b[o] = ab[o] * dactf(c->act_tanh, z2[o], y[o]);
        gb2[o] += w->z2b[o];
        real *gr = gW2 + (size_t)o * H;
        const real zb = w->z2b[o];
#pragma omp simd
        for (int jh = 0; jh < H; ++jh) gr[jh] += zb * h[jh];
    }
    for (int jh = 0; jh < H; ++jh) {
        real acc = 0;
        for (int o = 0; o < O; ++o) acc += W2[(size_t)o * H + jh] * w->z2b[o];
        w->z1b[jh] = acc * dactf(c->act_tanh, z1[jh], h[jh]);
        gb1[jh] += w->z1b[jh];
    }
    real *inpb = w->inpb;
    for (int i = 0; i < I; ++i) inpb[i] = 0;
    for (int jh = 0; jh < H; ++jh) {
        const real zb = w->z1b[jh];
        const real *wr = W1 + (size_t)jh * I;
        real *gr = gW1 + (size_t)jh * I;
#pragma omp simd
        for (int i = 0; i < I; ++i) {
            gr[i] += zb * inp[i];
            inpb[i] += zb * wr[i];
        }
    }
    /* sensor adjoint: COM velocity slots, vertex offsets (x_i - c) / scale */
    const real inv = (real)1 / (real)N;
    int q = c->n_sin;
    real vcb[3] = {inpb[q], inpb[q + 1], inpb[q + 2]};
    q += 3;
    real cb[3] = {0, 0, 0};
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            const real ob = inpb[q + 3 * i + d] / c->offset_scale;
            xb[3 * i + d] += ob;
            cb[d] -= ob;
        }
    for (int i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            xb[3 * i + d] += cb[d] * inv;
            vb[3 * i + d] += vcb[d] * inv;
        }
}

/* one task: forward with checkpoints, loss, backward with segment recompute */
static real task(const cr_cfg *c, work_t *w, const real *x0, const real tv[3], real th,
                 const real *W1, const real *b1, const real *W2, const real *b2, real scale,
                 real *g) {
    const int N3 = 3 * c->N, T = c->T, K = c->K, Wn = c->W;
    memcpy(w->x, x0, sizeof(real) * N3);
    memset(w->v, 0, sizeof(real) * N3);
    real total = 0;
    for (int t = 0; t <= T; ++t) {
        if (t % K == 0 && t < T) {
            memcpy(w->ck + (size_t)(t / K) * 2 * N3, w->x, sizeof(real) * N3);
            memcpy(w->ck + (size_t)(t / K) * 2 * N3 + N3, w->v, sizeof(real) * N3);
        }
        if (t > T - Wn) {
            real cm[3], vc[3];
            means(c, w->x, w->v, cm, vc);
            total += loss_term(cm, vc, tv, th);
        }
        if (t == T) break;
        step_fwd(c, w, 0, t, w->x, w->v, tv, th, W1, b1, W2, b2, w->xn, w->vn);
        real *tmp = w->x; w->x = w->xn; w->xn = tmp;
        tmp = w->v; w->v = w->vn; w->vn = tmp;
    }
    /* adjoint of the final state: its loss term */
    memset(w->xb1, 0, sizeof(real) * N3);
    memset(w->vb1, 0, sizeof(real) * N3);
    if (T > T - Wn) loss_adj(c, w->x, w->v, tv, th, scale, w->xb1, w->vb1);
    const int nseg = (T + K - 1) / K;
    for (int s = nseg - 1; s >= 0; --s) {
        const int t0 = s * K, n = T - t0 < K ? T - t0 : K;
        memcpy(w->xs, w->ck + (size_t)s * 2 * N3, sizeof(real) * N3);
        memcpy(w->vs, w->ck + (size_t)s * 2 * N3 + N3, sizeof(real) * N3);
        for (int j = 0; j < n; ++j) {
            real *xo = j + 1 < n ? w->xs + (size_t)(j + 1) * N3 : w->xn;
            real *vo = j + 1 < n ? w->vs + (size_t)(j + 1) * N3 : w->vn;
            step_fwd(c, w, j, t0 + j, w->xs + (size_t)j * N3, w->vs + (size_t)j * N3, tv, th,
                     W1, b1, W2, b2, xo, vo);
        }
        for (int j = n - 1; j >= 0; --j) {
            const int t = t0 + j;
            const real *x = w->xs + (size_t)j * N3, *v = w->vs + (size_t)j * N3;
            step_adj(c, w, j, t, x, v, W1, W2, g);
            if (t > T - Wn) loss_adj(c, x, v, tv, th, scale, w->xb, w->vb);
            real *tmp = w->xb1; w->xb1 = w->xb; w->xb = tmp;
            tmp = w->vb1; w->vb1 = w->vb; w->vb = tmp;
        }
    }
    return total / Wn;
}

int32_t cr_param_count(int32_t I, int32_t H, int32_t O) { return H * I + H + O * H + O; }

/* B tasks with initial positions x0 [B][N][3] (velocities zero), targets
 * tv [B][3], th [B]; writes loss_b [B] and grad [P] = d(mean loss)/dw in the
 * order W1, b1, W2, b2.  threads <= 0: OpenMP's default.  Returns 0, or -1
 * when out of memory. */
int32_t cr_rollout(const cr_cfg *c, int32_t B, const real *x0, const real *tv, const real *th,
                   const real *W1, const real *b1, const real *W2, const real *b2, real *loss_b,
                   real *grad, int32_t threads) {
    const int P = cr_param_count(c->I, c->H, c->O);
    real *gt = calloc((size_t)B * P, sizeof(real));
    if (!gt) return -1;
    int failed = 0;
    const real scale = (real)1 / ((real)c->W * (real)B);
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel num_threads(nt) reduction(| : failed)
    {
        work_t w;
        if (!work_alloc(&w, c)) {
            failed = 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int k = 0; k < B; ++k)
                loss_b[k] = task(c, &w, x0 + (size_t)k * 3 * c->N, tv + 3 * (size_t)k, th[k], W1,
                                 b1, W2, b2, scale, gt + (size_t)k * P);
        }
        work_free(&w);
    }
    if (!failed) {
        /* fixed-order sum over tasks (the result does not depend on threads) */
        for (int p = 0; p < P; ++p) grad[p] = 0;
        for (int k = 0; k < B; ++k)
            for (int p = 0; p < P; ++p) grad[p] += gt[(size_t)k * P + p];
    }
    free(gt);
    return failed ? -1 : 0;
}

/* Adam as the reference's Model.adam_update (beta1 = beta2 = 0.9 there), in
 * place; it = the 0-based iteration (bias corrections 1 - beta^(it + 1)) */
void cr_adam(int32_t n, real *w, const real *g, real *m, real *v, int32_t it, real lr, real b1,
             real b2, real eps) {
    const real c1 = (real)1 - r_pow(b1, it + 1), c2 = (real)1 - r_pow(b2, it + 1);
    for (int i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + ((real)1 - b1) * g[i];
        v[i] = b2 * v[i] + ((real)1 - b2) * g[i] * g[i];
        const real mh = m[i] / c1, vh = v[i] / c2;
        w[i] -= lr * mh / (r_sqrt(vh) + eps);
    }
}
