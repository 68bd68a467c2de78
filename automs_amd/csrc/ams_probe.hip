// ams_probe.hip -- VALU issue-rate probe of the measurement tooling.
//
// MI355X publishes no VALU instruction-issue rate, only FLOP rates (157.3
// TFLOP/s FP32 with packed FMA).  The tracer kernels are bound by VALU
// instruction issue, so their roofline needs that ceiling: these kernels
// issue streams of independent VALU instructions at full occupancy (eight
// chains per lane, 32 workgroups of 256 lanes per CU) for a rocprofv3 counter
// run (SQ_INSTS_VALU per CU and GRBM_GUI_ACTIVE cycle).  FP32 and 32-bit
// integer adds issue at close to two wave64 instructions per CU per cycle,
// FMA at one.
//   kind 0: non-packed FP32 FMA (v_fma_f32)
//   kind 1: 32-bit integer shift / xor / add
//   kind 2: FP32 add (v_add_f32)
//   kind 3: 32-bit integer add (v_add_u32)
// Every lane writes one float of out (blocks * 256 floats).
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ams.h"
#include "ams_internal.h"

namespace {

__global__ void __launch_bounds__(256) k_probe_fma(float *__restrict__ out, int iters) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = threadIdx.x * 0.001f + k;
  for (int i = 0; i < iters; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = fmaf(a[k], 0.99999f, 0.5f);
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_probe_int(float *__restrict__ out, int iters) {
  unsigned a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = threadIdx.x + k * 77u;
  for (int i = 0; i < iters; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = (a[k] ^ (a[k] >> 3)) + 0x9e3779b9u;
  unsigned s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
}

__global__ void __launch_bounds__(256) k_probe_addf(float *__restrict__ out, int iters) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = threadIdx.x * 0.001f + k;
  for (int i = 0; i < iters; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = a[k] + 0.5f;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_probe_addu(float *__restrict__ out, int iters) {
  unsigned a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = threadIdx.x + k;
  for (int i = 0; i < iters; i++)
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = a[k] + 0x9e3779b9u;
  unsigned s = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) s += a[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
}

}  // namespace

extern "C" int ams_valu_probe(float *out, int32_t blocks, int32_t iters, int32_t kind, void *stream) {
  if (!out || blocks <= 0 || iters <= 0 || kind < 0 || kind > 3) {
    ams::set_error("ams_valu_probe: bad arguments");
    return AMS_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  if (kind == 0) k_probe_fma<<<blocks, 256, 0, st>>>(out, iters);
  else if (kind == 1) k_probe_int<<<blocks, 256, 0, st>>>(out, iters);
  else if (kind == 2) k_probe_addf<<<blocks, 256, 0, st>>>(out, iters);
  else k_probe_addu<<<blocks, 256, 0, st>>>(out, iters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ams::set_error(std::string("ams_valu_probe: ") + hipGetErrorString(e));
    return AMS_ERR_HIP;
  }
  return AMS_OK;
}
