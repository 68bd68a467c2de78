/*
 * ams.h -- C ABI of the MI355X (gfx950) AutoMS channel tracer and phase-map
 * contraction.  Plain pointers and sizes only; no torch types.
 *
 * Every entry point returns AMS_OK (0) or a negative status; the message of
 * the last failure on the calling thread is returned by ams_last_error().
 * The reference catches and prints every error inside its op and returns
 * normally (optixPathTracer.cpp:858-860); these entry points report instead.
 */
#ifndef AMS_H_
#define AMS_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMS_OK 0
#define AMS_ERR_ARG (-1)
#define AMS_ERR_HIP (-2)
#define AMS_ERR_ALLOC (-3)
#define AMS_ERR_OVERFLOW (-4)

typedef struct ams_scene ams_scene;

/* A Tx or Rx element array.  Replaces the (src|dst, *_aabb_buffer, *_dir,
 * *_bptype, *_r) arguments of torch_launch_pathtrace
 * (optixPathTracer.cpp:781-791; layout [n0, n1, 3], n = n0*n1 at :814-821). */
typedef struct {
  int32_t n0, n1;   /* grid shape; element k = (k / n1, k % n1) */
  const float *pos; /* HOST pointer, [n0*n1][3] f32 */
  float dir[3];     /* pointing direction for the beam pattern */
  int32_t bptype;   /* beam pattern 0..3, optixPathTracer.cu:107-136 */
  float radius;     /* reception-sphere radius (go.py:633-634 defaults) */
} ams_array;

/* Trace options (the scalar arguments of torch_launch_pathtrace). */
typedef struct {
  double wave_len;        /* wave_len */
  int32_t max_reflection; /* max_reflection, 1..5 (MAX_HIT_N) */
  int32_t gridsize;       /* dst_girdsize: Rx tile edge */
  int64_t fib_n;          /* size of the Fibonacci direction set */
  int32_t per_element;    /* 1: fib_n rays from every Tx element; 0: fib_n rays in total,
                             ray r leaving element r % n_tx */
  int32_t flags;          /* bit 0: keep the reference's row wrap of the last tile column;
                             bit 1: evaluate the round-1 tasks only (no neighbour expansion);
                             bit 2: wavefront-persistent discovery (ray refill) instead of
                                    one thread per ray;
                             bit 3: after an Rx hit, query the scene again instead of keeping
                                    the scene hit already found on the same line;
                             bit 4: read the scene hierarchy from global memory even when it is
                                    small enough to be staged in LDS (32 KB);
                             bit 5: find Rx sphere hits through the sphere hierarchy also when
                                    the Rx array is a regular lattice;
                             bit 8: no analytic pre-check of backward chains (every chain is
                                    traced leg by leg);
                             bit 9: sort record keys at 128 bits even when they pack into 64;
                             bits 10, 11: forward queries walk the binary threaded hierarchy
                                     (bit 10) or the four-wide one (bit 11); by default the
                                     binary one for scenes staged in LDS and the four-wide one
                                     otherwise (same hits either way).  Bit 11 also makes the
                                     backward queries of scenes in global memory walk the
                                     four-wide hierarchy (by default they walk the binary one);
                             bit 7: keep tracing a branch with no reflection left even when it
                                    can no longer reach the Rx array;
                             bits 12, 13: build the neighbour-expansion rounds on the host
                                     (bit 12) or with kernels that expand over the traced-pair
                                     bitmap, group the new tasks by tile and form the units, the
                                     host reading two counts per round (bit 13, the default; same
                                     tasks and results either way); the task export and maps
                                     beyond 2^27 (path, tile) pairs use the host loop;
                             bits 14, 15: backward queries without beam candidate lists
                                     (bit 14: every leg walks the hierarchy) or with them for
                                     any Tx array (bit 15); by default the lists are built for
                                     Tx arrays of 256 elements and more.  A list holds every
                                     triangle within a margin of the convex beam that one leg
                                     of one (path, Rx) pair sweeps over all Tx elements, so a
                                     lane's closest hit among the candidates is the walk's
                                     (same results either way);
                             bit 16: with the lists, read a wave's lists and candidate
                                     triangles per lane even when every wave's lanes share one
                                     Rx (by default such reads are wave-uniform scalar loads
                                     when the Tx count is a multiple of 64; same results);
                             bit 18: a discovery branch with no reflection left queries the
                                     scene along the whole ray (by default only where the ray is
                                     in the Rx lattice's slab: elsewhere its pass-through hits
                                     cannot change the records; same records) */
  int32_t tasks_per_unit; /* tasks one workgroup loops over (0 = default 16) */
  void *stream;           /* hipStream_t; NULL = default stream */
  uint64_t *keys_out;     /* optional HOST buffer [keys_cap][2]: unique (path, tile) keys */
  int64_t keys_cap;
  const uint64_t *seed_keys; /* optional HOST [n_seed][2] (path, tile) keys used instead of
                                forward discovery */
  int64_t n_seed;
  int32_t discovery_mode;     /* 0: Fibonacci rays (fib_n, per_element); 1: the reference's
                                 rays: launch_x x launch_y threads x samples_per_launch, TEA/LCG
                                 random Tx and stratified directions, per-thread min_hc gating
                                 and reserve_size records (optixPathTracer.cu:142-190) */
  int32_t launch_x, launch_y, samples_per_launch, reserve_size;
  const char *task_info_path; /* optional: write one ams_task_info per backward chain of every
                                 expansion round to this file (the reference's path.dat
                                 under EXPORT_TASKINFO, optixPathTracer.cpp:244-256, :831) */
} ams_trace_opts;

/* One backward chain (optixPathTracer.h:40-52, 52 bytes in that layout).
 * Records of a round are ordered [src][task][Rx in tile] as the reference's
 * globalTaskId (optixPathTracer.cu:199), rounds one after another.  Fields
 * follow optixPathTracer.cu:243-250, :378-383, :407-412:
 *   successful     the chain reached its Tx element;
 *   receive_angle  cos(initial direction, dst_dir);
 *   launch_angle   cos(-final direction, src_dir), set on success;
 *   loss           the free-space factors only: each Fraunhofer split's
 *                  lambda / (4 pi D) and the final lambda / (4 pi D) at the Tx;
 *   phase          2 pi (path length) / lambda minus the reflection phases,
 *                  unwrapped, set on success;
 *   hit_cnt, hit_loss_idx  scene hits on the way, theta_bin + 19 * material.
 * Entries with no Rx element (tiles at the array's edge) and chains the
 * analytic pre-check skips hold zeros apart from src, dst and receive_angle
 * (the reference leaves the former unwritten). */
typedef struct {
  uint8_t successful;
  uint8_t pad_[3];
  float launch_angle, receive_angle, loss, phase;
  int32_t src, dst, hit_cnt;
  int32_t hit_loss_idx[5];
} ams_task_info;

/* *_nodes count 32-byte node boxes tested (a four-wide node is four) */
typedef struct {
  int64_t rays, records_raw, records_unique, paths, tasks, rounds, chains, successes, overflow;
  int64_t fwd_queries, fwd_nodes, fwd_tris, fwd_spheres;
  int64_t bwd_queries, bwd_nodes, bwd_tris;
  int64_t n_discover_launches, n_evaluate_launches;
  double t_discover_ms, t_evaluate_ms, t_total_ms; /* kernel times from HIP events; wall total */
  /* backward queries answered by a walk of the hierarchy (the others from
   * their beam's candidate list); beam lists: (task, Rx) pairs given lists,
   * their node and triangle tests, legs left to the walk, candidates listed */
  int64_t bwd_walks, beam_pairs, beam_nodes, beam_tris, beam_walk_legs, beam_cands;
} ams_trace_stats;

/* Scene: scene_triangle_buffer [3T,3] f32, mat [T] i32, mat_prop [n_mat,19]
 * f32 (optixPathTracer.cpp:808-813; go.py:679-737), all HOST pointers.
 * refl_phase [n_mat] (rad) may be NULL (reference: no reflection phase).
 * Builds the BVH once; the handle is reused across traces. */
int ams_scene_create(const float *tri, int32_t n_tri, const int32_t *mat, int32_t n_mat,
                     const float *mat_prop, const float *refl_phase, ams_scene **out);
int ams_scene_destroy(ams_scene *scene);
/* plane id of every triangle (smallest index of a coplanar triangle) */
int ams_scene_plane_ids(const ams_scene *scene, int32_t *out);

/* Replaces torch_launch_pathtrace (optixPathTracer.cpp:781-861): traces
 * Tx -> Rx and ACCUMULATES H = sum A e^{-j phi} into H_out, a DEVICE f32
 * buffer [n_tx][n_rx][2] (go.py:671-677, optixPathTracer.cu:259-264).
 * A scene handle serves one trace at a time (its device scratch and
 * page-locked staging buffers belong to the handle); traces that run
 * concurrently, on their own streams and host threads, use one handle
 * each.  Different handles share no state. */
int ams_trace_channel(ams_scene *scene, const ams_array *tx, const ams_array *rx,
                      const ams_trace_opts *opts, float *H_out, ams_trace_stats *stats);

/* ---- phase-map contraction (go.py:107-164) -----------------------------
 * Complex matrices are interleaved f32 (re, im) in DEVICE memory.
 * Fragment layout: a complex matrix X[R][Kd] split into bf16 hi + lo parts in
 * the lane order of v_mfma_f32_16x16x32_bf16, ams_frag_bytes(R, Kd) bytes. */
int64_t ams_frag_bytes(int32_t R, int32_t Kd);
/* pack X with element (r, k) at X + 2*(r*sr + k*sk) floats */
int ams_frag_pack(const float *X, int32_t R, int32_t Kd, int64_t sr, int64_t sk, void *frag, void *stream);
/* C[Ra][Rb] = A . B^T (conj_b bit 0: A . conj(B)^T), split-bf16 MFMA, FP32
 * accumulation; splits > 1 splits the reduction and needs
 * ams_cgemm_workspace_bytes(Ra, Rb, splits) of workspace.  Ra >= 128 uses
 * the LDS-staged workgroup kernel unless conj_b bit 1 asks for the per-wave
 * kernel (kept for comparison; identical results up to FP32 rounding).
 * With 8 or more splits the workgroup kernel gives each XCD whole splits, so
 * a split's slice of A crosses HBM once; conj_b bit 2 keeps the split-minor
 * order (identical results).  conj_b bit 3: the workgroup kernel forms each complex
 * tile product from three real products (9 MFMAs per tile pair instead of
 * 12, six planes per block staged in LDS; within 1e-5 of the four-product
 * form; measured slower on dense operands).  conj_b bit 4: the workgroup
 * kernel without its MFMA / LDS-read interleaving (identical results). */
int64_t ams_cgemm_workspace_bytes(int32_t Ra, int32_t Rb, int32_t splits);
int ams_cgemm_frag(const void *A, const void *B, int32_t Ra, int32_t Rb, int32_t Kd, int32_t conj_b, float *C,
                   int64_t ldc, int32_t splits, void *workspace, void *stream);
/* ams_cgemm_frag without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.  The
 * split partials live in a library-owned device array of
 * ams_cgemm_noalloc_max_bytes() (32 MiB) per GPU, held until the process
 * ends; a call whose ams_cgemm_workspace_bytes(Ra, Rb, splits) exceeds it
 * returns AMS_ERR_ARG.  The call allocates nothing.  Calls on one GPU share
 * that array: calls that overlap on one GPU (other streams or threads) race
 * and must be ordered by the caller. */
int64_t ams_cgemm_noalloc_max_bytes(void);
int ams_cgemm_frag_noalloc(const void *A, const void *B, int32_t Ra, int32_t Rb, int32_t Kd, int32_t conj_b,
                           float *C, int64_t ldc, int32_t splits, void *stream);
/* the same product in plain FP32 on the vector ALU (row strides in elements) */
int ams_cgemm_f32(const float *A, int64_t lda, const float *B, int64_t ldb, int32_t Ra, int32_t Rb, int32_t Kd,
                  int32_t conj_b, float *C, int64_t ldc, void *stream);

/* ---- coverage objective (go.py:159-245) and Adam (go.py:296-336) -------
 * Y complex f32 [G][C][N]: G groups (phase maps) x C codes x N Rx.
 * mode 0: L_g = -(1/N) sum_r log2(1 + max_c |y|^2 / n0)
 * mode 1: the reference capacity loss with row masks rowmask[G][C] (u8,
 *         device; group g masks rows rowmask[g * C + c]) and weight w
 *         (go.py:214-225). */
int64_t ams_objective_workspace_bytes(int32_t G, int32_t N);
int ams_objective_fwd(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                      const unsigned char *rowmask, float *loss, void *workspace, void *stream);
/* ams_objective_fwd with a coverage side output: the best code's received
 * power per Rx in dB (go.py:166-181, max over codes), averaged over Rx:
 * cov[g] = (1/N) sum_r 10 log10(max_c |y|^2 + 1e-30), over all codes
 * whatever the row mask; cov may be null. */
int ams_objective_fwd_cov(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                          const unsigned char *rowmask, float *loss, float *cov, void *workspace, void *stream);
int ams_objective_bwd(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                      const unsigned char *rowmask, const float *gloss, const void *workspace, float *gY,
                      void *stream);
/* The same gradient written directly in the fragment layout of
 * ams_frag_pack for the backward contraction: rows the G*C codes, reduction
 * index the N Rx (gY_frag: ams_frag_bytes(G*C, N) bytes); equal to
 * ams_frag_pack of ams_objective_bwd's gY. */
int ams_objective_bwd_frag(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                           const unsigned char *rowmask, const float *gloss, const void *workspace, void *gY_frag,
                           void *stream);

/* Phase-map operand of one metasurface with single-bounce chains, fused
 * (go.py:107-163 for one metasurface and max_ms_reflex = 1, replacing its
 * torch ops W = A e^{-j theta}, u = e^{j rem(phi)}, (W G) diag(u), cat).
 * Rows r = m*Cn + c (map m, code c) of A [nm*Cn][K+M] complex f32:
 *   A[r][k] = (W G)[c][k] u[m][k] (k < K),  A[r][K+j] = W[c][j],
 *   W[c][j] = amp[c][j] e^{-i th_ap[c][j]},
 *   u[m][k] = ms_amp[k] e^{i rem(th_ms[m][k], max_phase)} (torch.remainder).
 * th_ap, amp f32 [Cn][M]; th_ms f32 [nm][K]; ms_amp f32 [K]; G complex f32
 * [M][K].  Also writes WG = W G complex [Cn][K] and u complex [nm][K] for the
 * backward.  M <= 4096. */
int ams_pm_operand(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp, const float *G,
                   int32_t Cn, int32_t M, int32_t K, int32_t nm, float max_phase, float *A, float *WG, float *U,
                   void *stream);
/* The same operand written directly in the fragment layout of
 * ams_frag_pack (frag: ams_frag_bytes(nm*Cn, K+M) bytes), for
 * ams_cgemm_frag; equal to ams_frag_pack of ams_pm_operand's A.  WG and u as
 * there.  M <= 128. */
int ams_pm_operand_frag(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp,
                        const float *G, int32_t Cn, int32_t M, int32_t K, int32_t nm, float max_phase, void *frag,
                        float *WG, float *U, void *stream);
/* Its backward: gA complex [nm*Cn][K+M] in torch's convention for complex
 * gradients; g_ms f32 [nm][K] = d loss / d th_ms; when g_ap is not null also
 * g_ap f32 [Cn][M] = d loss / d th_ap, using gWG (complex [Cn][K]) as
 * workspace. */
int ams_pm_operand_bwd(const float *gA, const float *th_ap, const float *amp, const float *G, const float *WG,
                       const float *U, int32_t Cn, int32_t M, int32_t K, int32_t nm, float *g_ms, float *g_ap,
                       float *gWG, void *stream);
/* Metasurface chains over several surfaces (replaces the chain loop of
 * go.py:107-157, max_ms_reflex >= 1).  The surfaces' elements are
 * concatenated: K = sum of their element counts, th_ms [nm][K], ms_amp [K],
 * period [K] (the max phase of each element's surface), G [M][K] (the
 * src_to_ms channels side by side).  Rows r = m*Cn + c as above.
 * ams_chain_first: the first round, A [nm*Cn][K+M] = [WG o u | W] (the same
 * arithmetic as ams_pm_operand), WG [Cn][K], U [nm][K].
 * ams_chain_step: one later round after the caller's product Q = P . T
 * (ams_cgemm_frag against the [K][K] ms-to-ms block matrix T, zero diagonal
 * blocks): P = Q o u (P may be null in the last round), A[:, :K] += P (A at
 * row stride lda complex elements). */
int ams_chain_first(const float *th_ap, const float *amp, const float *th_ms, const float *ms_amp,
                    const float *period, const float *G, int32_t Cn, int32_t M, int32_t K, int32_t nm, float *A,
                    float *WG, float *U, void *stream);
int ams_chain_step(const float *Q, const float *U, int32_t Cn, int32_t nm, int32_t K, float *A, int64_t lda,
                   float *P, void *stream);
/* Backward of the chains.  gA [nm*Cn] rows at stride lda (torch's gradient
 * convention for complex tensors); Gacc [nm*Cn][K] (may be null) the part of
 * round r's output gradient that comes from the later rounds, so that
 * gP = gA + Gacc.  ams_chain_bwd_scale: gQ = gP o conj(u), the operand of the
 * next product gP_{r-1} - gA = gQ . conj(T)^T.  ams_chain_bwd_ms: the round's
 * share of the metasurface gradient, gu [nm][K] (+)= sum_c gP conj(X), X = Q_r
 * [nm*Cn][K] (x_by_code 0) or, in the first round, WG [Cn][K] (x_by_code 1);
 * g_ms (may be null) = Im(conj(u) gu) [nm][K] after the update, gWG (may be
 * null) [Cn][K] = sum_m gP conj(u).  ams_chain_bwd_ap: g_ap [Cn][M] from gWG
 * and gA's direct columns K..K+M-1 (lda = K + M), as ams_pm_operand_bwd. */
int ams_chain_bwd_scale(const float *gA, int64_t lda, const float *Gacc, const float *U, int32_t Cn, int32_t nm,
                        int32_t K, float *gQ, void *stream);
int ams_chain_bwd_ms(const float *gA, int64_t lda, const float *Gacc, const float *X, int32_t x_by_code,
                     const float *U, int32_t Cn, int32_t nm, int32_t K, float *gu, int32_t accumulate, float *g_ms,
                     float *gWG, void *stream);
int ams_chain_bwd_ap(const float *gA, const float *gWG, const float *G, const float *th_ap, const float *amp,
                     int32_t Cn, int32_t M, int32_t K, int32_t nm, float *g_ap, void *stream);
/* torch.optim.Adam step (no weight decay), step counted from 1 */
int ams_adam_step(float *param, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                  float beta2, float eps, int32_t step, void *stream);
/* The same step with its count kept on the device, and optionally the
 * best-loss tracking of the training loop (go.py:320-327, torch.optim.Adam's
 * state['step']).  step_state: AMS_ADAM_STATE_BYTES of device memory, zeroed
 * before the first step (its first int32 is the number of steps taken).
 * loss (device f32 scalar, may be null): the loss of param BEFORE this
 * update; when it is below every loss passed so far (the first one always,
 * unless NaN), param is copied to best_param (device, n floats; may be null)
 * before the update.  loss_hist (device, hist_cap floats; may be null)
 * receives loss at index (steps taken before this one) while below hist_cap.
 * Each call advances the state on the stream, so the call holds no host value
 * that changes between steps and can be captured in a HIP graph and replayed.
 * The parameters follow ams_adam_step's arithmetic. */
#define AMS_ADAM_STATE_BYTES 32
int ams_adam_step_dev(float *param, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                      float beta2, float eps, void *step_state, const float *loss, float *best_param,
                      float *loss_hist, int64_t hist_cap, void *stream);

/* ---- placement surrogate network (3x3 convolutions, bf16 weights) --------
 * X [B][Cin][H][W] f32, W [Cout][Cin][3][3] bf16, bias [Cout] f32, padding 1;
 * channels <= 32, H and W multiples of 16.  The input gradient is
 * ams_conv3x3_fwd with weights flipped in (ky, kx) and Cin/Cout swapped. */
int ams_conv3x3_fwd(const float *X, const void *Wbf16, const float *bias, float *Y, int32_t B, int32_t Cin,
                    int32_t Cout, int32_t H, int32_t W, int32_t relu, void *stream);
int ams_conv3x3_bwd_weight(const float *X, const float *dY, float *dW, float *db, int32_t B, int32_t Cin,
                           int32_t Cout, int32_t H, int32_t W, void *stream);
/* grad *= (Y > 0) */
/* The same convolution passes on bf16 MFMA (implicit GEMMs; activations and
 * output gradients split exactly into three bf16 parts, so the products
 * carry FP32 precision).  The
 * weights first go into the A-operand fragment layout,
 * ams_conv3x3_wfrag_bytes(Cout, Cin) bytes (for the input gradient: the
 * flipped, channel-transposed weights with Cin and Cout swapped).  The weight
 * gradient needs ams_conv3x3_dw_workspace_bytes(B, Cin, Cout, H, W) of
 * workspace for its per-tile partial sums, added in a fixed order.  B <=
 * 65535, channels <= 32, H and W multiples of 16. */
int64_t ams_conv3x3_wfrag_bytes(int32_t Cout, int32_t Cin);
int ams_conv3x3_pack_weights(const void *Wbf16, int32_t Cout, int32_t Cin, void *frag, void *stream);
int ams_conv3x3_fwd_mfma(const float *X, const void *Wfrag, const float *bias, float *Y, int32_t B, int32_t Cin,
                         int32_t Cout, int32_t H, int32_t W, int32_t relu, void *stream);
int64_t ams_conv3x3_dw_workspace_bytes(int32_t B, int32_t Cin, int32_t Cout, int32_t H, int32_t W);
int ams_conv3x3_bwd_weight_mfma(const float *X, const float *dY, float *dW, float *db, int32_t B, int32_t Cin,
                                int32_t Cout, int32_t H, int32_t W, void *workspace, void *stream);
int ams_relu_bwd(const float *Y, float *grad, int64_t n, void *stream);
/* mean((pred - target)^2 mask): grad [n] and per-block partial sums
 * [ams_mse_partials(n)] of the squared errors (sum them and divide by n) */
int64_t ams_mse_partials(int64_t n);
int ams_mse(const float *pred, const float *target, const float *mask, int64_t n, float *grad, float *partial,
            void *stream);
int ams_to_bf16(const float *x, void *y, int64_t n, void *stream);

/* ---- host-only helpers (no device calls; used by the CPU test suite) ---- */
/* the binned-SAH threaded BVH of ams_scene_create: nodes_out gets 32-byte
 * nodes {lo[3] f32, miss i32, hi[3] f32, info i32}, order_out the leaf order */
int ams_host_bvh_build(const double *lo, const double *hi, int32_t n, int32_t leaf_max, void *nodes_out,
                       int32_t nodes_cap, int32_t *order_out, int32_t *n_nodes_out);
/* the four-wide hierarchy the kernels walk, collapsed from the same binary
 * tree: nodes4_out gets 128-byte nodes {lo x[4], lo y[4], lo z[4], hi x[4],
 * hi y[4], hi z[4] f32, child[4] i32, pad[4]}; child >= 0 is a node, < 0 a
 * leaf -1 - ((first << 3) | count) into order_out, INT32_MIN an unused slot.
 * No counterpart in the reference (its hierarchy is built inside OptiX). */
int ams_host_bvh4_build(const double *lo, const double *hi, int32_t n, int32_t leaf_max, void *nodes4_out,
                        int32_t nodes4_cap, int32_t *order_out, int32_t *n_nodes4_out);
/* plane id of every triangle as ams_scene_plane_ids reports it */
int ams_host_plane_ids(const float *tri, int32_t n_tri, int32_t *plane_out);

/* ---- measurement tooling --------------------------------------------------
 * VALU issue-rate probe, no counterpart in the reference: independent VALU
 * instruction streams at full occupancy (kind 0: non-packed FP32 FMA, 1:
 * 32-bit integer shift / xor / add, 2: FP32 add, 3: 32-bit integer add) for a
 * counter run that calibrates the VALU-issue ceiling of the tracer kernels'
 * roofline.  out: device, blocks * 256 floats (one per lane). */
int ams_valu_probe(float *out, int32_t blocks, int32_t iters, int32_t kind, void *stream);

const char *ams_last_error(void);
int ams_version(void);
/* device bytes currently held by the library (cudaMemManager.h:75-81) */
int64_t ams_mem_outstanding(void);

#ifdef __cplusplus
}
#endif
#endif /* AMS_H_ */
