// peaks.hip -- issue rates the roofline figures are set against, measured.
//
// VALU: how many wave64 VALU instructions a gfx950 CU retires per cycle depends on
// the instruction: two-operand FP32/integer forms with VGPR operands (v_fmac_f32,
// v_add_f32, v_mul_f32, v_xor_b32, v_mov_b32) issue at close to two per CU per
// cycle; the same operations with an SGPR operand, three-operand forms, and
// v_cndmask/v_cmp/v_max/v_fma_mix at about one; the transcendental unit (v_exp,
// v_rsq, v_sqrt, v_rcp) at one half.  k_valu_form measures the rate of one form
// (v_fmac_f32 with VGPR operands among them, and the slower encodings the
// neighbour loops are made of) with independent chains at full occupancy.
//
// The bake's walk is a stream of unsigned min-atomics: ds_min_u32 into an LDS copy
// of a brick (the tiled schedules) and global atomicMin into the code grid (the
// plain schedule, and the tiled schedules' merge).  No datasheet gives their rates
// for gfx950, so these two kernels measure them, and the splat's atomics per
// second are reported as a fraction of the measured figure.
//
//   k_lds_min_peak     every lane min-merges into its own LDS word, 64 consecutive
//                      words per wave: no bank conflicts, the best case of the
//                      walk's LDS traffic.
//   k_global_min_peak  every lane min-merges into its own word of a buffer, a
//                      wave's 64 lanes on one 256-byte span, successive rounds
//                      striding over the whole buffer (far larger than L2).
//
// Both run a fixed grid for a bounded number of rounds; every address is inside
// the allocation by construction (lane index modulo the array size).
#include "clay_internal.h"

namespace clay {

namespace {

constexpr int kPeakThreads = 512;
constexpr int kLdsWords = 8192;  // 32 KB per workgroup

__global__ void __launch_bounds__(kPeakThreads) k_lds_min_peak(uint32_t rounds, uint32_t* __restrict__ sink) {
  __shared__ uint32_t lds[kLdsWords];
  const uint32_t t = threadIdx.x;
  for (uint32_t v = t; v < (uint32_t)kLdsWords; v += kPeakThreads) lds[v] = 0xFFFFFFFFu;
  __syncthreads();
  // Each round moves the whole workgroup to the next 512-word window: consecutive
  // lanes, consecutive words, every bank once per 32 lanes.
  uint32_t x = t * 2654435761u + blockIdx.x;
  for (uint32_t r = 0; r < rounds; r++) {
    const uint32_t w = (t + r * kPeakThreads) & (kLdsWords - 1u);
    x = x * 1664525u + 1013904223u;
    atomicMin(&lds[w], x);
  }
  __syncthreads();
  if (t == 0) sink[blockIdx.x] = lds[blockIdx.x & (kLdsWords - 1u)];
}

__global__ void __launch_bounds__(kPeakThreads)
    k_global_min_peak(uint32_t rounds, uint32_t* __restrict__ buf, uint32_t words_mask) {
  const uint32_t g = blockIdx.x * kPeakThreads + threadIdx.x;
  const uint32_t stride = gridDim.x * kPeakThreads;
  uint32_t x = g * 2654435761u;
  for (uint32_t r = 0; r < rounds; r++) {
    x = x * 1664525u + 1013904223u;
    atomicMin(&buf[(g + r * stride) & words_mask], x);
  }
}

// One instruction form per kernel, written out per step so the compiler can neither
// fuse nor re-encode it: eight independent register chains per lane, all
// register-to-register (no memory access inside the loop), and the result stored
// so nothing is folded away.
#define CLAY_RATE_OP(form, text, ...)                                                    \
  if constexpr (kForm == form) {                                                       \
    asm volatile(text : "+v"(a0) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a1) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a2) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a3) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a4) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a5) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a6) : __VA_ARGS__);                                       \
    asm volatile(text : "+v"(a7) : __VA_ARGS__);                                       \
  }

template <int kForm>
__global__ void __launch_bounds__(256) k_valu_form(uint32_t rounds, float xs, float ys, float* __restrict__ sink) {
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  const float x = xs * (1.0f + (float)(g & 3) * 0x1p-20f), y = ys * (1.0f + (float)(g & 7) * 0x1p-20f);
  const uint32_t h2 = 0x3c003c00u ^ (g & 1u);  // two f16 values near 1
  const uint64_t m = 0x5555555555555555ull;
  float a0 = 1e-3f * (float)g, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f, a4 = a0 + 4.0f, a5 = a0 + 5.0f,
        a6 = a0 + 6.0f, a7 = a0 + 7.0f;
  for (uint32_t r = 0; r < rounds; r++) {
    CLAY_RATE_OP(CLAY_VALU_FMAC, "v_fmac_f32_e32 %0, %1, %2", "v"(x), "v"(y))
    CLAY_RATE_OP(CLAY_VALU_FMAC_SGPR, "v_fmac_f32_e32 %0, %1, %2", "s"(xs), "v"(y))
    CLAY_RATE_OP(CLAY_VALU_FMA_VOP3, "v_fma_f32 %0, %0, %1, %2", "v"(x), "v"(y))
    CLAY_RATE_OP(CLAY_VALU_FMA_MIX, "v_fma_mix_f32 %0, %1, -1.0, %0 op_sel_hi:[1,0,0]", "v"(h2))
    CLAY_RATE_OP(CLAY_VALU_MAX, "v_max_f32_e32 %0, %0, %1", "v"(x))
    CLAY_RATE_OP(CLAY_VALU_CNDMASK, "v_cndmask_b32_e64 %0, %0, %1, %2", "v"(x), "s"(m))
    CLAY_RATE_OP(CLAY_VALU_RSQ, "v_rsq_f32_e32 %0, %0", "v"(x))
    CLAY_RATE_OP(CLAY_VALU_ADD_U32, "v_add_u32_e32 %0, %0, %1", "v"(x))
    CLAY_RATE_OP(CLAY_VALU_MUL_F32, "v_mul_f32_e32 %0, %0, %1", "v"(x))
  }
  sink[g] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}
#undef CLAY_RATE_OP

}  // namespace

int valu_peak(Ctx* c, int form, uint32_t rounds, double* wave_insts_per_s) {
  const unsigned blocks = (unsigned)c->cus * 32u;  // 32 four-wave groups per CU: several full occupancy rounds
  float* sink = nullptr;
  int rc = dalloc(&sink, (size_t)blocks * 256);
  if (rc) return rc;
  hipStream_t s = c->stream;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  float ms = 0.0f;
  for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // the first run warms up
    e = hipEventRecord(a, s);
    const float x = 0.999f, y = 1e-3f;
    switch (form) {
#define CLAY_RATE_CASE(f, kern) \
  case f: hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, rounds, x, y, sink); break;
      CLAY_RATE_CASE(CLAY_VALU_FMAC, k_valu_form<CLAY_VALU_FMAC>)
      CLAY_RATE_CASE(CLAY_VALU_FMAC_SGPR, k_valu_form<CLAY_VALU_FMAC_SGPR>)
      CLAY_RATE_CASE(CLAY_VALU_FMA_VOP3, k_valu_form<CLAY_VALU_FMA_VOP3>)
      CLAY_RATE_CASE(CLAY_VALU_FMA_MIX, k_valu_form<CLAY_VALU_FMA_MIX>)
      CLAY_RATE_CASE(CLAY_VALU_MAX, k_valu_form<CLAY_VALU_MAX>)
      CLAY_RATE_CASE(CLAY_VALU_CNDMASK, k_valu_form<CLAY_VALU_CNDMASK>)
      CLAY_RATE_CASE(CLAY_VALU_RSQ, k_valu_form<CLAY_VALU_RSQ>)
      CLAY_RATE_CASE(CLAY_VALU_ADD_U32, k_valu_form<CLAY_VALU_ADD_U32>)
      CLAY_RATE_CASE(CLAY_VALU_MUL_F32, k_valu_form<CLAY_VALU_MUL_F32>)
#undef CLAY_RATE_CASE
      default: break;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  dfree(sink);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("clay_valu_peak: ") + hipGetErrorString(e));
  *wave_insts_per_s = (double)blocks * (256 / 64) * 8.0 * rounds / (ms * 1e-3);
  return CLAY_OK;
}

int atomic_peak(Ctx* c, int kind, uint32_t rounds, double* ops_per_s) {
  const unsigned blocks = (unsigned)c->cus * 8u;
  const size_t words = (size_t)1 << 26;  // 256 MiB
  uint32_t* buf = nullptr;
  int rc = dalloc(&buf, kind == CLAY_ATOMIC_GLOBAL_MIN ? words : (size_t)blocks);
  if (rc) return rc;
  hipStream_t s = c->stream;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  float ms = 0.0f;
  for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {  // the first run warms up
    if (kind == CLAY_ATOMIC_GLOBAL_MIN)
      e = hipMemsetAsync(buf, 0xFF, words * sizeof(uint32_t), s);
    if (e == hipSuccess) e = hipEventRecord(a, s);
    if (kind == CLAY_ATOMIC_GLOBAL_MIN)
      hipLaunchKernelGGL(k_global_min_peak, dim3(blocks), dim3(kPeakThreads), 0, s, rounds, buf,
                         (uint32_t)(words - 1));
    else
      hipLaunchKernelGGL(k_lds_min_peak, dim3(blocks), dim3(kPeakThreads), 0, s, rounds, buf);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(b, s);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  dfree(buf);
  if (e != hipSuccess) return fail(CLAY_ERR_HIP, std::string("clay_atomic_peak: ") + hipGetErrorString(e));
  *ops_per_s = (double)blocks * kPeakThreads * rounds / (ms * 1e-3);
  return CLAY_OK;
}

}  // namespace clay

extern "C" int clay_atomic_peak(clay_ctx* ctx, int32_t kind, uint32_t rounds, double* ops_per_s) {
  clay::Ctx* c = reinterpret_cast<clay::Ctx*>(ctx);
  if (!c || !ops_per_s) return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: NULL argument");
  if (kind != CLAY_ATOMIC_LDS_MIN && kind != CLAY_ATOMIC_GLOBAL_MIN)
    return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: unknown atomic kind");
  if (rounds < 1 || rounds > 65536) return clay::fail(CLAY_ERR_INVALID, "clay_atomic_peak: rounds must be in [1, 65536]");
  return clay::atomic_peak(c, kind, rounds, ops_per_s);
}

extern "C" int clay_valu_peak(clay_ctx* ctx, int32_t form, uint32_t rounds, double* wave_insts_per_s) {
  clay::Ctx* c = reinterpret_cast<clay::Ctx*>(ctx);
  if (!c || !wave_insts_per_s) return clay::fail(CLAY_ERR_INVALID, "clay_valu_peak: NULL argument");
  if (form < CLAY_VALU_FMAC || form > CLAY_VALU_MUL_F32)
    return clay::fail(CLAY_ERR_INVALID, "clay_valu_peak: unknown instruction form");
  if (rounds < 1 || rounds > 1u << 20) return clay::fail(CLAY_ERR_INVALID, "clay_valu_peak: rounds must be in [1, 2^20]");
  return clay::valu_peak(c, form, rounds, wave_insts_per_s);
}
