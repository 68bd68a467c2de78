// ams_objective.hip -- coverage objectives over received signals y[g][c][r]
// (group g = one phase map, code c = one AP weight vector, r = Rx) and the
// fused Adam step of the phase-map optimiser.
//
// mode 0 (mean log SNR):  L_g = -(1/N) sum_r log2(1 + max_c |y_gcr|^2 / N0)
// mode 1 (reference capacity loss, go.py:166-237 with loss_mode "capacity"):
//   cap_cr = log2(1 + (|y_cr| + eps32)^2 / N0); rows with mask[c] = 1 are
//   zeroed (go.py:220); L = -w sum_r max_c cap' + || cap' . [cap' != max] ||_F
// Gradients follow PyTorch's convention for complex inputs,
// g = dL/dRe(y) + j dL/dIm(y).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/ams.h"
#include "ams_frag.h"
#include "ams_internal.h"

namespace {

constexpr float kEps32 = 1.1920928955078125e-07f;  // torch.finfo(torch.float32).eps
constexpr float kInvLn2 = 1.4426950408889634f;

__device__ __forceinline__ double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
  __syncthreads();
  return s;
}

__device__ __forceinline__ float cap_of(float2 y, float n0) {
  float a = sqrtf(y.x * y.x + y.y * y.y) + kEps32;
  return log2f(1.0f + a * a / n0);
}

// Coverage side output of one Rx: 10 log10 of the best code's power.
__device__ __forceinline__ double power_db(float p) { return (double)(10.0f * log10f(p + 1e-30f)); }

// Workgroup layout of the objective kernels: 64 Rx (one per lane) x 4 waves,
// wave q taking codes c = q, q+4, ...; per-Rx results of the four waves are
// combined through LDS.
constexpr int kObjRx = 64;
constexpr int kObjGroups = 4;

// The first (lowest-index) maximum over the four code groups of one Rx,
// each group's candidate being its own first maximum: the code a single
// ascending scan with a strict '>' would pick.
__device__ __forceinline__ void combine_first_max(const float (*val)[kObjRx], const int (*idx)[kObjRx], int rl,
                                                  float &best, int &bi) {
  best = val[0][rl];
  bi = idx[0][rl];
  for (int q = 1; q < kObjGroups; q++) {
    float v = val[q][rl];
    int i = idx[q][rl];
    if (v > best || (v == best && i < bi)) {
      best = v;
      bi = i;
    }
  }
}

// partial[g][block] = (S1, S2, S3): mode 0: S1 = sum log2(1+max SNR); mode 1:
// S1 = sum max cap', S2 = sum of squares of non-max cap'.  S3 (when
// want_cov) = sum of power_db(max_c |y|^2).
__global__ void __launch_bounds__(kObjRx *kObjGroups) k_obj_fwd(const float2 *__restrict__ Y, int C, int N,
                                                                float n0, int mode,
                                                                const unsigned char *__restrict__ rowmask,
                                                                int want_cov, double *__restrict__ partial,
                                                                int *__restrict__ argmax) {
  __shared__ double sh[8];
  __shared__ float val[kObjGroups][kObjRx];
  __shared__ int idx[kObjGroups][kObjRx];
  __shared__ float pw[kObjGroups][kObjRx];
  const int g = blockIdx.y;
  const int rl = threadIdx.x % kObjRx, q = threadIdx.x / kObjRx;
  const int r = blockIdx.x * kObjRx + rl;
  const bool ok = r < N;
  const float2 *Yg = Y + (size_t)g * C * N;
  const unsigned char *rm = rowmask ? rowmask + (size_t)g * C : nullptr;  // the group's row mask
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  // pass 1: each group's first maximum (power in mode 0, capacity in mode 1)
  float best = -1.0f, pmax = 0.0f;
  int bi = 0;
  if (ok)
    for (int c = q; c < C; c += kObjGroups) {
      float2 y = Yg[(size_t)c * N + r];
      float p = y.x * y.x + y.y * y.y;
      pmax = fmaxf(pmax, p);
      float v = mode == 0 ? p : (rm && rm[c] ? 0.0f : cap_of(y, n0));
      if (v > best) {
        best = v;
        bi = c;
      }
    }
  val[q][rl] = best;
  idx[q][rl] = bi;
  pw[q][rl] = pmax;
  __syncthreads();
  combine_first_max(val, idx, rl, best, bi);
  if (ok && q == 0) {
    if (mode == 0) {
      s1 = (double)log2f(1.0f + best / n0);
      if (want_cov) s3 = power_db(best);
    } else {
      float pm = fmaxf(fmaxf(pw[0][rl], pw[1][rl]), fmaxf(pw[2][rl], pw[3][rl]));
      s1 = best;
      if (want_cov) s3 = power_db(pm);
    }
    if (argmax) argmax[(size_t)g * N + r] = bi;
  }
  // pass 2 (mode 1): squares of the codes whose capacity is not the maximum
  if (mode != 0 && ok)
    for (int c = q; c < C; c += kObjGroups) {
      float v = rm && rm[c] ? 0.0f : cap_of(Yg[(size_t)c * N + r], n0);
      if (v != best) s2 += (double)v * v;
    }
  double t1 = block_sum(s1, sh);
  double t2 = mode != 0 ? block_sum(s2, sh) : 0.0;
  double t3 = want_cov ? block_sum(s3, sh) : 0.0;
  if (threadIdx.x == 0) {
    double *p = partial + 3 * ((size_t)g * gridDim.x + blockIdx.x);
    p[0] = t1;
    p[1] = t2;
    p[2] = t3;
  }
}

// loss[g] (and cov[g]) from the partials: lane l adds blocks l, l+64, ... in
// order, then a fixed shuffle tree (deterministic).
__global__ void __launch_bounds__(64) k_obj_finish(const double *__restrict__ partial, int nb, int N, int mode,
                                                   float w, double *__restrict__ sums, float *__restrict__ loss,
                                                   float *__restrict__ cov) {
  const int g = blockIdx.x;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int b = threadIdx.x; b < nb; b += 64) {
    const double *p = partial + 3 * ((size_t)g * nb + b);
    s1 += p[0];
    s2 += p[1];
    s3 += p[2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    s1 += __shfl_down(s1, off);
    s2 += __shfl_down(s2, off);
    s3 += __shfl_down(s3, off);
  }
  if (threadIdx.x != 0) return;
  sums[2 * g] = s1;
  sums[2 * g + 1] = s2;
  if (mode == 0) loss[g] = (float)(-s1 / N);
  else loss[g] = (float)(-(double)w * s1 + sqrt(s2));
  if (cov) cov[g] = (float)(s3 / N);
}

// d L / d y of one element (code c of a group, one Rx), in torch's convention
// for complex gradients.  bi: the Rx's argmax code; best: its maximum
// capacity (mode 1); fro: the Frobenius norm of the non-maximum capacities
// (mode 1); masked: the code's row is masked (mode 1).
__device__ __forceinline__ float2 obj_grad(float2 y, int c, int bi, float best, float fro, bool masked, int mode,
                                           float n0, float w, float gl, int N) {
  float2 out = make_float2(0.0f, 0.0f);
  if (mode == 0) {
    if (c == bi) {
      float s = y.x * y.x + y.y * y.y;
      // dL/ds = -(1/N) / (ln2 (N0 + s)); g = 2 y dL/ds
      float dls = -gl * kInvLn2 / ((float)N * (n0 + s));
      out = make_float2(2.0f * dls * y.x, 2.0f * dls * y.y);
    }
    return out;
  }
  if (masked) return out;
  float v = cap_of(y, n0);
  float dcap = 0.0f;
  if (c == bi) dcap -= w;
  if (v != best && fro > 0.0f) dcap += v / fro;
  float amp = sqrtf(y.x * y.x + y.y * y.y);
  float a = amp + kEps32;
  float dca = kInvLn2 * 2.0f * a / n0 / (1.0f + a * a / n0);  // dcap / d|y|
  float gabs = gl * dcap * dca;
  if (amp > 0.0f) out = make_float2(gabs * y.x / amp, gabs * y.y / amp);
  return out;
}

__global__ void __launch_bounds__(kObjRx *kObjGroups) k_obj_bwd(const float2 *__restrict__ Y, int C, int N,
                                                                float n0, int mode, float w,
                                                                const unsigned char *__restrict__ rowmask,
                                                                const int *__restrict__ argmax,
                                                                const double *__restrict__ sums,
                                                                const float *__restrict__ gloss,
                                                                float2 *__restrict__ gY) {
  __shared__ float val[kObjGroups][kObjRx];
  const int g = blockIdx.y;
  const int rl = threadIdx.x % kObjRx, q = threadIdx.x / kObjRx;
  const int r = blockIdx.x * kObjRx + rl;
  const bool ok = r < N;
  const float2 *Yg = Y + (size_t)g * C * N;
  float2 *Gg = gY + (size_t)g * C * N;
  const unsigned char *rm = rowmask ? rowmask + (size_t)g * C : nullptr;
  const float gl = gloss[g];
  const int bi = ok ? argmax[(size_t)g * N + r] : 0;
  if (mode == 0) {
    if (!ok) return;
    for (int c = q; c < C; c += kObjGroups)
      Gg[(size_t)c * N + r] =
          c == bi ? obj_grad(Yg[(size_t)c * N + r], c, bi, 0.0f, 0.0f, false, 0, n0, w, gl, N) : make_float2(0, 0);
    return;
  }
  // mode 1: the maximum capacity of the Rx (over all four groups) first
  float best = -1.0f;
  if (ok)
    for (int c = q; c < C; c += kObjGroups) {
      float v = rm && rm[c] ? 0.0f : cap_of(Yg[(size_t)c * N + r], n0);
      if (v > best) best = v;
    }
  val[q][rl] = best;
  __syncthreads();
  best = fmaxf(fmaxf(val[0][rl], val[1][rl]), fmaxf(val[2][rl], val[3][rl]));
  if (!ok) return;
  const float fro = (float)sqrt(sums[2 * g + 1]);
  for (int c = q; c < C; c += kObjGroups)
    Gg[(size_t)c * N + r] =
        obj_grad(Yg[(size_t)c * N + r], c, bi, best, fro, rm && rm[c], 1, n0, w, gl, N);
}

// The same gradient written straight into the split-bf16 fragment layout of
// the backward contraction (ams_frag.h; rows = the G*C codes, reduction
// index = Rx): one thread per fragment lane, 8 consecutive Rx of one code.
// The maximum capacity of an Rx (mode 1) is that of its argmax code.  Equal
// to ams_frag_pack of k_obj_bwd's output.
__global__ void __launch_bounds__(256) k_obj_bwd_frag(const float2 *__restrict__ Y, int G, int C, int N, float n0,
                                                      int mode, float w, const unsigned char *__restrict__ rowmask,
                                                      const int *__restrict__ argmax,
                                                      const double *__restrict__ sums,
                                                      const float *__restrict__ gloss, int n_kb, long long n_items,
                                                      u32x4 *__restrict__ frag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const ams::FragItem it = ams::frag_item(t, n_kb);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; e++) v[e] = make_float2(0.0f, 0.0f);
  if (it.r < (long long)G * C) {
    const int g = (int)(it.r / C), c = (int)(it.r % C);
    const float2 *Yg = Y + (size_t)g * C * N;
    const float gl = gloss[g];
    const float fro = mode != 0 ? (float)sqrt(sums[2 * g + 1]) : 0.0f;
    const unsigned char *rm = rowmask ? rowmask + (size_t)g * C : nullptr;
    const bool masked = mode != 0 && rm && rm[c];
    // the lane's 8 values of y and their argmax codes are loaded together (one
    // round of memory latency), then combined
    float2 yc[8];
    int bis[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int rx = it.k0 + e;
      yc[e] = rx < N ? Yg[(size_t)c * N + rx] : make_float2(0.0f, 0.0f);
      bis[e] = rx < N ? argmax[(size_t)g * N + rx] : -1;
    }
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int rx = it.k0 + e;
      if (rx >= N) continue;
      const int bi = bis[e];
      if (mode == 0) {
        if (c == bi) v[e] = obj_grad(yc[e], c, bi, 0.0f, 0.0f, false, 0, n0, w, gl, N);
      } else {
        const float best = rm && rm[bi] ? 0.0f : cap_of(Yg[(size_t)bi * N + rx], n0);
        v[e] = obj_grad(yc[e], c, bi, best, fro, masked, 1, n0, w, gl, N);
      }
    }
  }
  ams::frag_store8(frag, it, v);
}

// torch.optim.Adam (no weight decay, no amsgrad), single-tensor form:
// m = lerp(m, g, 1-b1);  v = b2 v + (1-b2) g^2;
// p -= (lr / (1-b1^t)) * m / (sqrt(v) / sqrt(1-b2^t) + eps)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, long long n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2_sqrt) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = m[i] + (1.0f - b1) * (gi - m[i]);
  float vi = v[i] * b2 + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float denom = sqrtf(vi) / bc2_sqrt + eps;
  p[i] = p[i] - (lr / bc1) * (mi / denom);
}

// Device-side state of the fused Adam (AMS_ADAM_STATE_BYTES): the step
// count, the bias corrections of the last step, and the best-loss tracking
// of go.py:320-327 (the loss of the parameters before the update is compared
// with the best so far; when lower, those parameters are kept).  done counts
// the workgroups of a launch that have finished; it is zero between launches.
struct AdamState {
  int32_t step;
  float bc1;
  float bc2_sqrt;
  int32_t better;
  float best_loss;
  int32_t best_valid;
  uint32_t done;
  int32_t pad;
};

// One launch per step.  Every lane reads the state the previous step left
// and derives this step's count, bias corrections (the arithmetic of the
// host path: double pow, rounded to float once) and best-loss decision
// itself.  The last workgroup to finish (found through done, which it
// resets) writes the new state and the history entry; every workgroup has
// read the old state before it counts itself done, so none can see the new.
__global__ void __launch_bounds__(256) k_adam_dev(float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, long long n,
                                                  float lr, float b1, float b2, float eps, AdamState *st,
                                                  const float *__restrict__ loss, float *__restrict__ best_p,
                                                  float *__restrict__ hist, long long hist_cap) {
  __shared__ int last;
  const int32_t t = st->step + 1;
  const float bc1 = (float)(1.0 - pow((double)b1, (double)t));
  const float bc2_sqrt = (float)sqrt(1.0 - pow((double)b2, (double)t));
  const float l = loss ? *loss : 0.0f;
  const float best = st->best_valid ? st->best_loss : INFINITY;
  const int better = loss && l < best;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float pi = p[i];
    if (best_p && better) best_p[i] = pi;
    float gi = g[i];
    float mi = m[i] + (1.0f - b1) * (gi - m[i]);
    float vi = v[i] * b2 + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float denom = sqrtf(vi) / bc2_sqrt + eps;
    p[i] = pi - (lr / bc1) * (mi / denom);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&st->done, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last || threadIdx.x != 0) return;
  st->done = 0;
  st->step = t;
  st->bc1 = bc1;
  st->bc2_sqrt = bc2_sqrt;
  st->better = better;
  if (better) {
    st->best_loss = l;
    st->best_valid = 1;
  }
  if (loss && hist && t - 1 < hist_cap) hist[t - 1] = l;
}

}  // namespace

extern "C" {

int64_t ams_objective_workspace_bytes(int32_t G, int32_t N) {
  int64_t nb = (N + kObjRx - 1) / kObjRx;
  return (int64_t)G * nb * 24 + (int64_t)G * 16 + (int64_t)G * N * 4;
}

// Y complex f32 [G][C][N]; loss f32 [G]; cov f32 [G] or null; workspace from
// ams_objective_workspace_bytes (keeps argmax and sums for the backward).
int ams_objective_fwd_cov(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                          const unsigned char *rowmask, float *loss, float *cov, void *workspace, void *stream) {
  if (!Y || !loss || !workspace || G <= 0 || C <= 0 || N <= 0 || (mode != 0 && mode != 1) || !(n0 > 0)) {
    ams::set_error("ams_objective_fwd: bad arguments");
    return AMS_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  int nb = (N + kObjRx - 1) / kObjRx;
  double *partial = (double *)workspace;
  double *sums = partial + 3 * (size_t)G * nb;
  int *argmax = (int *)(sums + 2 * (size_t)G);
  k_obj_fwd<<<dim3(nb, G), kObjRx * kObjGroups, 0, st>>>((const float2 *)Y, C, N, n0, mode, rowmask, cov ? 1 : 0, partial, argmax);
  k_obj_finish<<<G, 64, 0, st>>>(partial, nb, N, mode, w, sums, loss, cov);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_objective_fwd: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_objective_fwd(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                      const unsigned char *rowmask, float *loss, void *workspace, void *stream) {
  return ams_objective_fwd_cov(Y, G, C, N, mode, n0, w, rowmask, loss, nullptr, workspace, stream);
}

// gY complex f32 [G][C][N] = gloss[g] * dL_g/dY (workspace of the forward).
int ams_objective_bwd(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                      const unsigned char *rowmask, const float *gloss, const void *workspace, float *gY,
                      void *stream) {
  if (!Y || !gY || !gloss || !workspace || G <= 0 || C <= 0 || N <= 0 || (mode != 0 && mode != 1)) {
    ams::set_error("ams_objective_bwd: bad arguments");
    return AMS_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  int nb = (N + kObjRx - 1) / kObjRx;
  const double *partial = (const double *)workspace;
  const double *sums = partial + 3 * (size_t)G * nb;
  const int *argmax = (const int *)(sums + 2 * (size_t)G);
  k_obj_bwd<<<dim3(nb, G), kObjRx * kObjGroups, 0, st>>>((const float2 *)Y, C, N, n0, mode, w, rowmask, argmax, sums, gloss,
                                         (float2 *)gY);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_objective_bwd: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_objective_bwd_frag(const float *Y, int32_t G, int32_t C, int32_t N, int32_t mode, float n0, float w,
                           const unsigned char *rowmask, const float *gloss, const void *workspace, void *gY_frag,
                           void *stream) {
  if (!Y || !gY_frag || !gloss || !workspace || G <= 0 || C <= 0 || N <= 0 || (mode != 0 && mode != 1)) {
    ams::set_error("ams_objective_bwd_frag: bad arguments");
    return AMS_ERR_ARG;
  }
  int nb = (N + kObjRx - 1) / kObjRx;
  const double *partial = (const double *)workspace;
  const double *sums = partial + 3 * (size_t)G * nb;
  const int *argmax = (const int *)(sums + 2 * (size_t)G);
  const long long R = (long long)G * C;
  const int n_kb = (N + 31) / 32;
  const long long n_items = ((R + 15) / 16) * n_kb * 64;
  k_obj_bwd_frag<<<(unsigned)((n_items + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      (const float2 *)Y, G, C, N, n0, mode, w, rowmask, argmax, sums, gloss, n_kb, n_items, (u32x4 *)gY_frag);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_objective_bwd_frag: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_adam_step(float *param, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                  float beta2, float eps, int32_t step, void *stream) {
  if (!param || !grad || !m || !v || n <= 0 || step < 1) {
    ams::set_error("ams_adam_step: bad arguments");
    return AMS_ERR_ARG;
  }
  double bc1 = 1.0 - std::pow((double)beta1, step);
  double bc2 = 1.0 - std::pow((double)beta2, step);
  k_adam<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(param, grad, m, v, n, lr, beta1, beta2, eps,
                                                                       (float)bc1, (float)std::sqrt(bc2));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_adam_step: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

int ams_adam_step_dev(float *param, const float *grad, float *m, float *v, int64_t n, float lr, float beta1,
                      float beta2, float eps, void *step_state, const float *loss, float *best_param,
                      float *loss_hist, int64_t hist_cap, void *stream) {
  if (!param || !grad || !m || !v || !step_state || n <= 0 || (best_param && !loss) || (loss_hist && !loss) ||
      hist_cap < 0) {
    ams::set_error("ams_adam_step_dev: bad arguments");
    return AMS_ERR_ARG;
  }
  static_assert(sizeof(AdamState) == AMS_ADAM_STATE_BYTES, "AdamState layout");
  hipStream_t st = (hipStream_t)stream;
  AdamState *as = (AdamState *)step_state;
  k_adam_dev<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(param, grad, m, v, n, lr, beta1, beta2, eps, as, loss,
                                                          best_param, loss_hist, (long long)hist_cap);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_adam_step_dev: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}

}  // extern "C"
