/*
 * diffloco.h -- C ABI of the MI355X (gfx950) differentiable locomotion rollout.
 *
 * Every entry point is extern "C", noexcept, takes plain pointers and sizes,
 * and returns 0 on success or a negative DL_E* code; dl_last_error() returns
 * the message of the calling thread's last failure.  Device buffers are
 * caller-allocated (torch) device pointers; `stream` is a hipStream_t (NULL =
 * default stream).  Nothing here allocates device memory.  `dtype` selects
 * the arithmetic type of every real-valued buffer of the call (DL_F32 or
 * DL_F64); integer buffers are int32.
 *
 * Field layout of the per-step API (one design, B robots): positions and
 * velocities are SoA per robot and step, x[t][k][d][i] (d = 0..2, i = vertex);
 * step slices are passed as pointers to x[t].
 *
 * Each per-step entry point replaces one reference kernel (file:line in the
 * reference repository, erizmr/Complex-locomotion-skill-learning-via-
 * differentiable-physics); the *_grad entry points replace the adjoint that
 * Taichi generates for that kernel under ti.ad.Tape
 * (multitask/diffphy_trainer.py:246).  Gradients ACCUMULATE (+=) into the
 * *_grad outputs, like Taichi's .grad fields.
 */
#ifndef DIFFLOCO_H
#define DIFFLOCO_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DL_ABI_VERSION 1

enum { DL_F32 = 0, DL_F64 = 1 };
enum { DL_ACT_SIN = 0, DL_ACT_TANH = 1 };
enum { DL_OK = 0, DL_E_ARG = -1, DL_E_HIP = -2, DL_E_SIZE = -3, DL_E_NUMERIC = -4 };

/* Physics constants of the mass-spring step
 * (multitask/solver_mass_spring.py:8-23, multitask/config_sim.py:112-121). */
typedef struct dl_ms_physics {
    double dt;
    double gravity;          /* along +y */
    double ground;           /* ground height */
    double drag;             /* velocity factor exp(-dt*drag) (:120) */
    double dashpot;          /* dashpot impulse coefficient, no dt (:103) */
    double friction;         /* per-axis Coulomb clamp (:136-144) */
} dl_ms_physics;

/* Sensor vector layout (multitask/taichi_env.py:149-183, generalised):
 *   [n_sin sinusoids | vcom? 3 COM-velocity slots | per vertex (x-c)/offset_scale
 *    followed by vert_pad zeros | dup_v copies of target_v[tv_comp[..]]/max_speed
 *    | dup_h copies of target_h | dup_c copies of target_c] */
typedef struct dl_sensor_layout {
    int32_t n_sin;
    int32_t run_period;      /* sinusoid period in steps */
    int32_t vcom;            /* 0/1 */
    int32_t vert_pad;
    int32_t n_tv;            /* components per target_v copy (1..3) */
    int32_t tv_comp[3];
    int32_t dup_v, dup_h, dup_c;
    int32_t pad_;
    double offset_scale;
    double max_speed;
} dl_sensor_layout;

/* Topology of one design (multitask/solver_mass_spring.py:47-67): springs
 * (endpoints, rest length, stiffness ALREADY multiplied by the solver's
 * stiffness scale, actuation coefficient, driving MLP output) plus two CSR
 * maps the builder derives: vertex -> incident springs (entry = 2*s + side,
 * side 0 when the vertex is endpoint a, 1 when b) and MLP output -> springs. */
typedef struct dl_ms_topology {
    int32_t n_vert, n_spring, n_act;
    int32_t pad_;
    const int32_t *spring_a, *spring_b, *spring_act;    /* [S] */
    const void *spring_L0, *spring_k, *spring_coeff;    /* [S] real */
    const int32_t *vert_ptr, *vert_idx;                 /* [N+1], [2S] */
    const int32_t *act_ptr, *act_idx;                   /* [O+1], [<=S] */
} dl_ms_topology;

/* ---------------------------------------------------------------- generic */
int32_t dl_abi_version(void);
int32_t dl_last_error(char *buf, size_t len);
int32_t dl_device_count(int32_t *count);

/* ------------------------------------------------- per-step mass-spring API */

/* compute_center (solver_mass_spring.py:150-156): center[k] = mean_i x[k][:,i].
 * Also writes the COM velocity when vcom != NULL (quadruped sensor/loss). */
int32_t dl_ms_center(int32_t dtype, int32_t B, int32_t N, const void *x_t, const void *v_t,
                     void *center_t, void *vcom_t, void *stream);
int32_t dl_ms_center_grad(int32_t dtype, int32_t B, int32_t N, const void *center_grad,
                          const void *vcom_grad, void *x_grad, void *v_grad, void *stream);

/* compute_height (:158-173): height[t] = windowed max of min_i y, upper = max_i y.
 * height_prev may be NULL when t % jump_period == 0. */
int32_t dl_ms_height(int32_t dtype, int32_t B, int32_t N, int32_t t, int32_t jump_period,
                     const void *x_t, const void *height_prev, void *height_t, void *upper_t,
                     void *stream);
int32_t dl_ms_height_grad(int32_t dtype, int32_t B, int32_t N, int32_t t, int32_t jump_period,
                          const void *x_t, const void *height_prev, const void *height_grad,
                          const void *upper_grad, void *x_grad, void *height_prev_grad,
                          void *stream);

/* compute_rotation (:175-192).  robot_id 100 uses vertices 45,46,34,36;
 * otherwise head/tail split by x_0 < center_0 along x. */
int32_t dl_ms_rotation(int32_t dtype, int32_t B, int32_t N, int32_t robot_id, const void *x_t,
                       const void *x_0, const void *center_0, void *rotation_t, void *stream);
int32_t dl_ms_rotation_grad(int32_t dtype, int32_t B, int32_t N, int32_t robot_id,
                            const void *x_t, const void *x_0, const void *center_0,
                            const void *rotation_grad, void *x_grad, void *stream);

/* nn_input (taichi_env.py:149-183).  tv_t [B][3], th_t/tc_t [B]; in_t [B][I]. */
int32_t dl_nn_input(int32_t dtype, const dl_sensor_layout *lay, int32_t B, int32_t N,
                    int32_t t, int32_t offset, const void *x_t, const void *v_t,
                    const void *center_t, const void *tv_t, const void *th_t,
                    const void *tc_t, void *in_t, void *stream);
int32_t dl_nn_input_grad(int32_t dtype, const dl_sensor_layout *lay, int32_t B, int32_t N,
                         const void *in_grad, void *x_grad, void *v_grad, void *center_grad,
                         void *stream);

/* Model.nn1 + nn2 (nn.py:130-150): hidden = act(W1 in + b1), out = act(W2 hidden + b2).
 * W1 [H][I], W2 [O][H].  hidden_act [B][H], out_act [B][O]. */
int32_t dl_mlp_forward(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H, int32_t O,
                       const void *in, const void *W1, const void *b1, const void *W2,
                       const void *b2, void *hidden_act, void *out_act, void *stream);
/* Adjoint of dl_mlp_forward given out_grad; accumulates into in_grad (may be
 * NULL), W1_grad, b1_grad, W2_grad, b2_grad.  work: caller workspace of
 * B*(2H+O) reals.  Weight gradients are summed over the batch in a fixed
 * order (deterministic).  FP32 with H a multiple of 16, H <= 64, I <= 256
 * and O <= 16 runs the products X W1^T, z1b W1 and z1b^T X on the matrix
 * cores (MFMA 16x16x4 FP32 tiles of 16 robots); other shapes and FP64 run
 * the VALU kernels. */
int32_t dl_mlp_backward(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H, int32_t O,
                        const void *in, const void *W1, const void *b1, const void *W2,
                        const void *b2, const void *out_grad, void *in_grad, void *W1_grad,
                        void *b1_grad, void *W2_grad, void *b2_grad, void *work, void *stream);

/* Library options (process-wide).  DL_OPT_MLP_MFMA for the eligible FP32
 * calls of dl_mlp_forward / dl_mlp_backward: 1 (default) the faster kernel
 * of each stage (forward and weight gradients on MFMA, the input adjoint on
 * VALU), 2 every stage on MFMA, 0 the VALU kernels only (for comparison). */
enum { DL_OPT_MLP_MFMA = 1 };
int32_t dl_set_option(int32_t option, int32_t value);

/* dl_mlp_backward without the caller workspace: allocates nothing (no
 * hipMalloc, hipFree or allocating library call) and uses a library-owned
 * device array of DL_MLP_STATIC_WORK_BYTES bytes per GPU, held until the
 * process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Requires B*(2H+O)*sizeof(real) <=
 * DL_MLP_STATIC_WORK_BYTES, else DL_E_ARG.  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (different streams or host
 * threads) race and must be serialised by the caller.  Results equal
 * dl_mlp_backward's (same kernels). */
#define DL_MLP_STATIC_WORK_BYTES (4 << 20)
int32_t dl_mlp_backward_static(int32_t dtype, int32_t act, int32_t B, int32_t I, int32_t H,
                               int32_t O, const void *in, const void *W1, const void *b1,
                               const void *W2, const void *b2, const void *out_grad,
                               void *in_grad, void *W1_grad, void *b1_grad, void *W2_grad,
                               void *b2_grad, void *stream);

/* advance(t) = apply_spring_force(t) + advance_toi(t+1)
 * (solver_mass_spring.py:86-148, :200-202), fused: reads x[t], v[t], act[t]
 * ([B][O]), writes x[t+1], v[t+1].  Spring impulses are summed per vertex in
 * CSR order (deterministic). */
int32_t dl_ms_advance(int32_t dtype, const dl_ms_topology *topo, const dl_ms_physics *phys,
                      int32_t B, const void *x_t, const void *v_t, const void *act_t,
                      void *x_t1, void *v_t1, void *stream);
int32_t dl_ms_advance_grad(int32_t dtype, const dl_ms_topology *topo,
                           const dl_ms_physics *phys, int32_t B, const void *x_t,
                           const void *v_t, const void *act_t, const void *x_t1_grad,
                           const void *v_t1_grad, void *x_t_grad, void *v_t_grad,
                           void *act_t_grad, void *stream);

/* compute_loss_velocity / compute_loss_rotation (taichi_env.py:185-211, 235-252)
 * over a stored trajectory.  center [T][B][3], x [T][B][3][N], tv [T][B][3],
 * th [T][B].  The rollout wrote center rows t < T_written and x rows
 * t <= T_written (x also holds the state after the last advance); unwritten
 * rows read as zero.  loss_batch [B] receives loss*B per robot like
 * loss_velocity_batch; loss [1] the sum. */
int32_t dl_loss_velocity(int32_t dtype, int32_t B, int32_t T, int32_t T_written, int32_t steps,
                         int32_t run_period, int32_t turn_period, const void *center,
                         const void *tv, void *loss, void *loss_batch, void *stream);
int32_t dl_loss_velocity_grad(int32_t dtype, int32_t B, int32_t T, int32_t T_written,
                              int32_t steps, int32_t run_period, int32_t turn_period,
                              const void *center, const void *tv, double loss_grad,
                              void *center_grad, void *stream);
/* work: caller workspace of B reals. */
int32_t dl_loss_rotation(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t T_written,
                         int32_t steps, int32_t run_period, int32_t turn_period, const void *x,
                         const void *center, const void *th, void *loss, void *loss_batch,
                         void *work, void *stream);
int32_t dl_loss_rotation_grad(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t T_written,
                              int32_t steps, int32_t run_period, int32_t turn_period,
                              const void *x, const void *center, const void *th,
                              double loss_grad, void *x_grad, void *center_grad, void *stream);

/* compute_loss_height (taichi_env.py:213-222) over height [T][B], target_h
 * [T][B] (rows t <= steps < T read; height rows >= T_written read as zero):
 * sum over t in [1, steps] with t % jump_period == jump_period - 1 and
 * target_h > 0.1 + 1e-6 of (height - target_h)^2 / B / (steps / jump_period)
 * * 100.  loss_batch [B] = per-robot sum * B (may be NULL); max_height [1]
 * (may be NULL) is raised to the largest contributing height
 * (current_max_height).  work: B*2 reals. */
int32_t dl_loss_height(int32_t dtype, int32_t B, int32_t T, int32_t T_written, int32_t steps,
                       int32_t jump_period, const void *height, const void *target_h, void *loss,
                       void *loss_batch, void *max_height, void *work, void *stream);
int32_t dl_loss_height_grad(int32_t dtype, int32_t B, int32_t T, int32_t T_written,
                            int32_t steps, int32_t jump_period, const void *height,
                            const void *target_h, double loss_grad, void *height_grad,
                            void *stream);
/* compute_loss_pose (taichi_env.py:224-233): x [T][B][3][N], initial objects
 * [N][3]; sum over t in [1, steps], t % jump_period == 0, of |x - obj|^2 /
 * B / (steps / jump_period).  work: B reals. */
int32_t dl_loss_pose(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t T_written,
                     int32_t steps, int32_t jump_period, const void *x, const void *objects,
                     void *loss, void *loss_batch, void *work, void *stream);
int32_t dl_loss_pose_grad(int32_t dtype, int32_t B, int32_t N, int32_t T, int32_t T_written,
                          int32_t steps, int32_t jump_period, const void *x, const void *objects,
                          double loss_grad, void *x_grad, void *stream);
/* compute_loss_actuation (taichi_env.py:276-285): actuation [T][B][O],
 * target_v [T][B][3], target_h [T][B]; sum over t in [0, steps) with
 * target_h < 0.1 + 1e-4 of max(|a| - sqrt(|tv_x| / 0.08), 0) / O / B / steps
 * * 10 (O = the reference's spring count for its robots).  Adjoint: sign(a)
 * where |a| exceeds the allowance.  work: B reals. */
int32_t dl_loss_actuation(int32_t dtype, int32_t B, int32_t O, int32_t T, int32_t T_written,
                          int32_t steps, const void *act, const void *target_v,
                          const void *target_h, void *loss, void *loss_batch, void *work,
                          void *stream);
int32_t dl_loss_actuation_grad(int32_t dtype, int32_t B, int32_t O, int32_t T, int32_t T_written,
                               int32_t steps, const void *act, const void *target_v,
                               const void *target_h, double loss_grad, void *act_grad,
                               void *stream);
/* compute_loss_crawl (taichi_env.py:287-293): upper_height, target_c [T][B];
 * sum over t in [1, steps] with target_c > 1 - 1e-4 of max(upper - 0.1, 0)^2
 * / B / steps * 5.  work: B reals. */
int32_t dl_loss_crawl(int32_t dtype, int32_t B, int32_t T, int32_t T_written, int32_t steps,
                      const void *upper_height, const void *target_c, void *loss,
                      void *loss_batch, void *work, void *stream);
int32_t dl_loss_crawl_grad(int32_t dtype, int32_t B, int32_t T, int32_t T_written, int32_t steps,
                           const void *upper_height, const void *target_c, double loss_grad,
                           void *upper_grad, void *stream);

/* Windowed COM loss of the quadruped task over stored center/vcom [T+1][B][3]
 * with per-robot targets tv [B][3], th [B]:
 *   loss = mean_k mean_{t in (T-W, T]} |vc_xz - tv_xz|^2 + (c_y - th)^2.
 * No reference counterpart (the reference trains velocity/rotation losses,
 * taichi_env.py:185-252); it is the objective of the quadruped workload. */
int32_t dl_loss_task(int32_t dtype, int32_t B, int32_t T, int32_t W, const void *center,
                     const void *vcom, const void *tv, const void *th, void *loss,
                     void *loss_batch, void *stream);
int32_t dl_loss_task_grad(int32_t dtype, int32_t B, int32_t T, int32_t W, const void *center,
                          const void *vcom, const void *tv, const void *th, double loss_grad,
                          void *center_grad, void *vcom_grad, void *stream);

/* ------------------------------------------- persistent batched rollout API */

/* One row of the robot offset table (device memory).  Robots of different
 * size, design and horizon share one launch; one workgroup advances one
 * robot through many steps with its state resident in LDS. */
typedef struct dl_robot {
    int32_t n_vert, n_spring, n_act, n_in;
    int32_t horizon;         /* number of advances T */
    int32_t t_fwd;           /* forward: next step to take (0 initially) */
    int32_t t_bwd;           /* backward: adjoint has reached this step (T initially) */
    int32_t spring0;         /* first spring of the design in the spring arrays */
    int32_t vptr0;           /* first row of the design's vertex CSR (N+1 rows) */
    int32_t aptr0;           /* first row of the design's actuator CSR (O+1 rows) */
    int64_t weights;         /* element offset of the design's weight block */
    int64_t wgrad;           /* element offset of this robot's gradient partial */
    int64_t state;           /* element offset of the [6][N] state / adjoint slot */
    int64_t ckpt;            /* element offset of [T/K+1][6][N] checkpoints */
    int64_t scratch;         /* element offset of the per-segment recompute buffer */
    int64_t targets;         /* element offset of targets: [1 or T][5] = tv xyz, th, tc */
    int64_t traj;            /* element offset of [T][3N+3] x/center trajectory, -1 = none */
    int32_t target_stride;   /* 0 = constant per robot, 5 = per step */
    int32_t pad_;
} dl_robot;

/* Weight block of a design: W1 [H][I], W1T [I][H], b1 [H], W2 [O][H],
 * W2T [H][O], b2 [O] -- dl_weights_size() elements, rounded up to a
 * multiple of 16, so blocks packed back to back from an aligned base keep
 * W1 and W1T 64-byte aligned (FP32 robots with W1 outside LDS read W1T as
 * 16-byte rows when H = 64 and it is aligned).  Gradient partial
 * block: W1 [H][I], b1 [H], W2 [O][H], b2 [O] -- dl_wgrad_size() elements. */
int64_t dl_weights_size(int32_t I, int32_t H, int32_t O);
int64_t dl_wgrad_size(int32_t I, int32_t H, int32_t O);
/* Elements of per-robot scratch for checkpoint interval K: two buffers of K
 * steps' adjoint rows -- state (6N), per-vertex impulse sums (3N, written by
 * the stored-rows forward only), network input (I), hidden values,
 * derivatives and adjoints (3H), outputs (3O), COM row (6); the pipelined
 * adjoint fills one buffer while it reverses the other -- and a spring spill
 * area (10S + 4: the block kernels keep spilled spring rows as 16-byte rows,
 * 8 reals per spring, from an aligned start). */
int64_t dl_scratch_size(int32_t N, int32_t S, int32_t I, int32_t H, int32_t O, int32_t K);

typedef struct dl_rollout {
    int32_t dtype, act;
    int32_t n_robots;
    int32_t n_hidden;        /* H, shared by every design */
    int32_t ckpt_every;      /* K */
    int32_t max_steps;       /* steps this launch may take per robot (<= 0: to the end) */
    int32_t block;           /* threads per workgroup (64..512, multiple of 64; the
                                wavefront-per-robot variants take 64 or 256) */
    int32_t loss_kind;       /* 0: none in-kernel, 1: windowed COM loss */
    int32_t loss_window;     /* states t in (T - window, T] */
    int32_t store_traj;      /* forward writes x/center trajectory rows */
    int32_t max_n, max_s, max_i, max_o;   /* maxima over the table (LDS sizing) */
    int32_t lds_policy;      /* 0 auto, 1 stage topology, 2 topology + W1, 3 none,
                                4 as much as fits in the whole LDS */
    int32_t variant;         /* 0 auto, 1 workgroup per robot, 2 wavefront per robot (FP32,
                                H = 64, I <= 192, O <= 8; robots of one design packed by 4),
                                3 as 2 with the adjoint's segment recompute on a second
                                wavefront per robot, pipelined over two scratch buffers,
                                4 as 3 with no recompute: the forward writes every step's
                                adjoint rows into scratch buffer 0, which then needs
                                ckpt_every >= every horizon (a robot whose horizon
                                exceeds it gets a NaN loss / gradient),
                                5 as 1 with the forward's rows kept likewise */
    dl_ms_physics phys;
    dl_sensor_layout lay;
    const dl_robot *robots;  /* device [n_robots] */
    const int32_t *spring_a, *spring_b, *spring_act;
    const void *spring_L0, *spring_k, *spring_coeff;
    const int32_t *vert_ptr, *vert_idx, *act_ptr, *act_idx;
    const void *weights;
    void *wgrad;             /* per-robot gradient partials (accumulated) */
    void *state;             /* x,v slots (forward state) */
    void *adjoint;           /* xbar,vbar slots (backward state) */
    void *ckpt;
    void *scratch;
    const void *targets;
    void *loss;              /* [n_robots], accumulated by the forward */
    const void *loss_grad;   /* [n_robots] upstream dL/dloss_r for the backward */
    void *traj;              /* forward trajectory rows */
    const void *traj_grad;   /* backward: external dL/dx_t, dL/dc_t rows (NULL = none) */
    int32_t *err;            /* optional device word (NULL = none): the launches OR in
                                DL_ROLLOUT_ERR_* bits when a robot breaks the table's
                                contract (that robot's loss or gradient is also set to NaN
                                and it is skipped); zero it before the launches */
} dl_rollout;

enum { DL_ROLLOUT_ERR_HORIZON = 1,   /* stored rows (variant 4/5): horizon > ckpt_every */
       DL_ROLLOUT_ERR_PACKING = 2 }; /* wave kernels: robots of one workgroup of 4 do not
                                        share a design */

/* Advance every robot of the table from its t_fwd by up to max_steps steps
 * (one persistent launch), writing checkpoints every K steps. */
int32_t dl_ms_rollout_forward(const dl_rollout *r, void *stream);
/* Adjoint of the rollout: from each robot's t_bwd down by up to max_steps
 * steps, segment by segment (recompute from checkpoint, then reverse). */
int32_t dl_ms_rollout_backward(const dl_rollout *r, void *stream);
/* Sum per-robot gradient partials into per-design gradients in a fixed order:
 * out[d] = sum over robots r with design_of[r] == d of wgrad[r]. */
int32_t dl_wgrad_reduce(int32_t dtype, int32_t n_robots, const int32_t *robot_list_ptr,
                        const int32_t *robot_list, int32_t n_designs,
                        const int64_t *robot_wgrad_off, const int64_t *design_out_off,
                        const int64_t *design_size, const void *wgrad, void *out,
                        void *stream);

/* ------------------------------------------------------------- MLS-MPM API */

/* MLS-MPM constants (multitask/solver_mpm.py:23-33, config_sim.py:112-129).
 * Active stress A = act_strength * a * e e^T along axis act_axis (reference:
 * y in 2D). */
typedef struct dl_mpm_params {
    int32_t dim;             /* 2 or 3 */
    int32_t n_grid;          /* n cells per axis, dx = 1/n */
    int32_t bound;           /* sticky-wall width in cells */
    int32_t act_axis;
    double dt, gravity, p_vol, p_mass, mu, la, act_strength;
} dl_mpm_params;

/* Particle fields per robot, structure-of-arrays: x, v [B][D][P]; C, F
 * [B][D*D][P] (row-major entries); act [B][n_act]; act_id [B][P] (-1 =
 * passive).  Grids per robot: grid_v [B][D][n^D], grid_m [B][n^D]. */

/* Fill n reals with a value (grid clears; graph-capture safe). */
int32_t dl_mpm_fill(int32_t dtype, void *ptr, int64_t n, double value, void *stream);
/* p2g (solver_mpm.py:83-124): F update, polar, stress, APIC scatter into the
 * (cleared) grid with atomics; writes F_out.  *err |= 1 if a stencil leaves
 * the grid. */
int32_t dl_mpm_p2g(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P,
                   const void *x, const void *v, const void *C, const void *F, const void *act,
                   const int32_t *act_id, int32_t n_act, void *F_out, void *grid_v,
                   void *grid_m, int32_t *err, void *stream);
/* grid_op (solver_mpm.py:126-144). */
int32_t dl_mpm_grid(int32_t dtype, const dl_mpm_params *params, int32_t B, const void *grid_v,
                    const void *grid_m, void *grid_v_out, void *stream);
/* g2p (solver_mpm.py:146-165). */
int32_t dl_mpm_g2p(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P,
                   const void *x, const void *grid_v_out, void *x_out, void *v_out, void *C_out,
                   void *stream);
/* Adjoints (the Taichi-generated g2p.grad, grid_op.grad, p2g.grad that
 * advance_core_grad runs, solver_mpm.py:182-190).  g2p_adj WRITES x_grad
 * (= x_grad1 + stencil terms) and accumulates grid_v_out_grad atomically;
 * grid_adj writes the grid adjoints; p2g_adj accumulates x_grad, WRITES
 * v_grad, C_grad, F_grad, and accumulates act_grad atomically. */
int32_t dl_mpm_g2p_adj(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P,
                       const void *x, const void *grid_v_out, const void *x_grad1,
                       const void *v_grad1, const void *C_grad1, void *x_grad,
                       void *grid_v_out_grad, void *stream);
int32_t dl_mpm_grid_adj(int32_t dtype, const dl_mpm_params *params, int32_t B,
                        const void *grid_v, const void *grid_m, const void *grid_v_out_grad,
                        void *grid_v_grad, void *grid_m_grad, void *stream);
int32_t dl_mpm_p2g_adj(int32_t dtype, const dl_mpm_params *params, int32_t B, int32_t P,
                       const void *x, const void *v, const void *C, const void *F,
                       const void *act, const int32_t *act_id, int32_t n_act,
                       const void *grid_v_grad, const void *grid_m_grad, const void *F_grad1,
                       void *x_grad, void *v_grad, void *C_grad, void *F_grad, void *act_grad,
                       void *stream);
/* Sensor vector of an MPM robot (3D): [n_sin sinusoids | COM velocity |
 * per-group mean offset from the COM / offset_scale | target_v / max_speed |
 * target_h]; targets [B][5]; group [B][P] in [0, n_groups); inv_count [B][G]
 * = 1 / particles per group.  Also writes center/vcom [B][3] when non-NULL.
 * No reference counterpart: the reference's MPM env path is unreachable
 * (taichi_env.py:136 calls SolverMPM() without its config). */
int32_t dl_mpm_sensor(int32_t dtype, int32_t B, int32_t P, int32_t n_groups, int32_t n_sin,
                      int32_t run_period, int32_t t, double offset_scale, double max_speed,
                      const void *x, const void *v, const int32_t *group, const void *targets,
                      const void *inv_count, void *in, void *center, void *vcom, void *stream);
int32_t dl_mpm_sensor_adj(int32_t dtype, int32_t B, int32_t P, int32_t n_groups, int32_t n_sin,
                          double offset_scale, const int32_t *group, const void *inv_count,
                          const void *in_grad, void *x_grad, void *v_grad, void *stream);

/* ---------------------------------------------------------------- batched MPM rollout
 * B robots of one particle design, each on its own grid, advanced (and
 * differentiated) a segment of steps per call.  Replaces the per-step loop
 * of SolverMPM.advance_core / advance_core_grad (solver_mpm.py:175-190) plus
 * the controller (nn.py:130-150) and COM/sensor reductions
 * (solver_mpm.py:167-173) of a training rollout, with the Taichi tape's
 * recompute scheme (checkpoint -> recompute the segment -> adjoint).
 *
 * Grids: [B][n^D][4] records.  grid = (momentum (D), mass), grid_vo =
 * velocity, grid_vo_grad its adjoint, grid_grad = adjoint of grid.  Active
 * blocks (4^3 nodes in 3D, 8^2 in 2D): blk_flag, blk_list [B][(n/E)^D],
 * blk_count [B].  Between calls grid, grid_vo_grad, blk_flag, blk_count and
 * act_grad must be zero (allocate them zeroed; the calls keep them so).
 * Per-step rows (index t = 0..horizon): sens [T+1][B][3*n_groups+6] (sums:
 * per-group x, x, v; zeroed by the caller before a forward rollout),
 * center/vcom [T+1][B][3], inp [T][B][I], hid [T][B][H], act_out [T][B][O],
 * in_grad [T+1][B][I] (row T zero), center_grad/vcom_grad [T+1][B][3] (loss
 * adjoints, e.g. from dl_loss_task_grad), wgrad [B][H*I + H + O*H + O]
 * (per-robot slices, accumulated).  I = n_sin + 3 + 3*n_groups + 4. */
enum { DL_MPM_CONTROL = 1, DL_MPM_SENSE_FIRST = 2, DL_MPM_SAVE_GRID = 4,
       DL_MPM_SAVE_HORIZON = 8, DL_MPM_REPLAY = 16 };
/* Scatter tiles: particles are processed DL_MPM_PB per workgroup
 * (W = ceil(P / DL_MPM_PB) workgroups per robot), each accumulating its
 * stencil box in an LDS tile of tile_nodes nodes (a box larger than that is
 * done in slabs).  scatter = DL_MPM_SCATTER_ATOMIC commits tiles to the grid
 * with global atomics; DL_MPM_SCATTER_STORE stores them to tiles
 * [B][W][tile_nodes][4] (box in tile_hdr [B][W][8]), lists them under every
 * block they overlap (blk_tile_count [B][nblk], zero between calls;
 * blk_tiles [B][nblk][DL_MPM_TILE_CAP]) and the grid pass gathers them. */
enum { DL_MPM_PB = 256, DL_MPM_TILE_MAX = 8192, DL_MPM_TILE_CAP = 32 };
enum { DL_MPM_SCATTER_ATOMIC = 0, DL_MPM_SCATTER_STORE = 1 };

typedef struct dl_mpm_batch {
    int32_t dtype, B, P, n_act;
    int32_t n_groups;        /* sensor groups (actuators + passive body) */
    int32_t n_sin, run_period, n_hidden;
    int32_t act;             /* DL_ACT_* of both MLP layers */
    int32_t horizon;         /* T */
    int32_t tile_nodes;      /* LDS scatter tile capacity, nodes */
    int32_t scatter;         /* DL_MPM_SCATTER_* */
    double offset_scale, max_speed;
    dl_mpm_params params;
    const int32_t *act_id;   /* [B][P], -1 = passive */
    const int32_t *group;    /* [B][P] in [0, n_groups) */
    const void *inv_count;   /* [B][n_groups] */
    const void *targets;     /* [B][5]: tv (3), th, unused */
    const void *W1, *b1, *W2, *b2;
    void *grid, *grid_vo, *grid_vo_grad, *grid_grad;
    int32_t *blk_flag, *blk_list, *blk_count;
    int32_t *err;            /* |= 1 when a particle's stencil leaves the grid */
    void *sens, *center, *vcom, *inp, *hid, *act_out;
    void *act_grad;          /* [B][O] */
    void *in_grad;           /* NULL: no controller adjoint (physics only) */
    void *center_grad, *vcom_grad, *wgrad;
    void *tiles;
    int32_t *tile_hdr, *blk_tile_count, *blk_tiles;
    /* Saved grids of a segment (optional, all NULL or none): for segment
     * step j, the active blocks' grid records and velocities grid_save /
     * vo_save [K][B][nblk][E^D][4] (by list position), list_save [K][B][nblk],
     * count_save [K][B].  dl_mpm_forward with DL_MPM_SAVE_GRID fills them;
     * dl_mpm_backward then restores each step's grid instead of re-running
     * P2G and the grid update (same values, one pass fewer). */
    void *grid_save, *vo_save;
    int32_t *list_save, *count_save;
    int32_t save_steps;      /* K: steps the save buffers hold */
    int32_t scatter_unroll;  /* stencil-offset loop unroll of the scatter kernels: 1, 3, 9, 27 */
    /* Controller adjoint rows (required with in_grad): hid_grad [T][B][H] and
     * out_grad [T][B][O], the adjoints of each step's hidden and output
     * pre-activations.  dl_mpm_backward writes rows t0..t0+n-1 and adds their
     * sums over those steps (outer products with the inp / hid rows) to
     * wgrad once per call. */
    void *hid_grad, *out_grad;
    /* Scatter windows (optional): windows [n_windows][2] = (first, end)
     * particle index ranges that partition [0, P) in order, each of at most
     * DL_MPM_PB particles.  P2G and the G2P adjoint run one workgroup per
     * window (NULL: ceil(P / DL_MPM_PB) windows of DL_MPM_PB particles), so
     * cutting the particle order where a window's stencil box would grow
     * large keeps every workgroup's box small.  Store-mode tile buffers hold
     * one tile per window.  A window wider than DL_MPM_PB sets err bit 2. */
    const int32_t *windows;
    int32_t n_windows, pad_;
    /* Horizon grid save (optional): dl_mpm_forward with DL_MPM_SAVE_HORIZON
     * stores every step's active blocks, at most hcap per robot, in
     * hgrid / hvo [T][B][hcap][64][4] (totals, velocities), hlist [T][B][hcap]
     * and hcount [T][B]; hover [T] is set (nonzero) for a step where some robot
     * had more active blocks (that robot's step is not saved).  A segment
     * whose steps are all saved may be recomputed with DL_MPM_REPLAY: per step
     * the saved velocities are restored and only the update of F and G2P run
     * (no P2G, no grid update; x, v, C come out as in the forward).
     * dl_mpm_backward with restore_from = 1 restores each step's grid from
     * this save (0: from grid_save, or by P2G when that is NULL). */
    void *hgrid, *hvo;
    int32_t *hlist, *hcount, *hover;
    int32_t hcap, restore_from;
    /* Uneven batches (optional; NULL / 0: every robot has P particles, the
     * whole horizon and the shared weights).  n_part [B]: robot k's particles
     * are slots [0, n_part[k]) of its P (the rest are padding, never read;
     * sensor and COM sums divide by n_part[k]).  horizons [B]: robot k
     * advances steps t < horizons[k] only (its state t = horizons[k] is its
     * last; the controller, P2G, grid and G2P kernels and their adjoints
     * skip the robot at later steps; the caller seeds its loss gradient at
     * its own steps).  per_robot_weights = 1: W1, b1, W2, b2 hold B blocks
     * (robot k's at k*H*I, k*H, k*O*H, k*O reals).  Uneven batches run with
     * plain windows (windows = NULL) and without the horizon save. */
    const int32_t *n_part, *horizons;
    int32_t per_robot_weights, pad2_;
} dl_mpm_batch;

/* Ring of particle-state slots: slot s of x/v at x + s*B*D*P reals, of C/F
 * at C + s*B*D*D*P reals. */
typedef struct dl_mpm_slots {
    void *x, *v, *C, *F;
    int32_t n_slots, pad_;
} dl_mpm_slots;

/* Steps t0 .. t0+n-1: state t0 in slot 0, state t0+j in slot j % n_slots.
 * DL_MPM_CONTROL: run the controller (center/vcom/inp/hid/act_out rows t,
 * sensor sums of row t+1; center/vcom of row T when t0+n == T); without it
 * act_out rows are read (segment recompute).  DL_MPM_SENSE_FIRST: first
 * accumulate the sensor sums of state t0 (row t0).  DL_MPM_SAVE_GRID /
 * DL_MPM_SAVE_HORIZON: save each step's grid (segment slots / horizon rows;
 * not both).  DL_MPM_REPLAY: recompute from the horizon save (no
 * DL_MPM_CONTROL, no saving). */
int32_t dl_mpm_forward(const dl_mpm_batch *b, const dl_mpm_slots *states, int32_t t0,
                       int32_t n, int32_t flags, void *stream);
/* Adjoint of steps t0 .. t0+n-1 given the segment's states (n+1 slots) and
 * the adjoint of state t0+n in slot 0 of `adj` (2 slots), without the
 * sensor/COM terms of that state, which this call adds from in_grad,
 * center_grad and vcom_grad rows t0+n.  On return slot 0 holds the adjoint
 * of state t0 in the same convention; in_grad rows t0..t0+n-1 and wgrad are
 * updated. */
int32_t dl_mpm_backward(const dl_mpm_batch *b, const dl_mpm_slots *segment,
                        const dl_mpm_slots *adj, int32_t t0, int32_t n, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* DIFFLOCO_H */
