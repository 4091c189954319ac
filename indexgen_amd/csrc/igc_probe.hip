// igc_probe.hip -- VALU issue-rate probe: the int32 instruction types of the
// pair kernel (xor, shift, find-first-bit, max3, or3, add3, add), each in 8
// independent chains per thread (inline ALU asm so nothing folds), one store
// per thread.  Gives the practical VALU peak the pair kernel is measured
// against.
#include "igc_device.hpp"

namespace igc {

constexpr int kProbeIters = 2048;

#define IGC_PROBE_ST(ins)                                                                                     \
    asm volatile(ins : "+v"(a0) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a1) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a2) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a3) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a4) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a5) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a6) : "s"(k1));                                                                   \
    asm volatile(ins : "+v"(a7) : "s"(k1));
#define IGC_PROBE_4(ins) IGC_PROBE_ST(ins) IGC_PROBE_ST(ins) IGC_PROBE_ST(ins) IGC_PROBE_ST(ins)

template <int OP>
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t *__restrict__ out, uint32_t seed)
{
    uint32_t a0 = seed ^ threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3, a4 = a0 + 4, a5 = a0 + 5, a6 = a0 + 6,
             a7 = a0 + 7;
    const uint32_t k1 = __builtin_amdgcn_readfirstlane(seed * 0x9E37u | 1u);
    for (int i = 0; i < kProbeIters; ++i) {
        if constexpr (OP == 0) { IGC_PROBE_4("v_xor_b32 %0, %1, %0") }
        if constexpr (OP == 1) { IGC_PROBE_4("v_lshrrev_b32 %0, %1, %0") }
        if constexpr (OP == 2) { IGC_PROBE_4("v_ffbl_b32 %0, %0") }
        if constexpr (OP == 3) { IGC_PROBE_4("v_max3_u32 %0, %0, %1, %0") }
        if constexpr (OP == 4) { IGC_PROBE_4("v_or3_b32 %0, %0, %1, %0") }
        if constexpr (OP == 5) { IGC_PROBE_4("v_add3_u32 %0, %0, %1, %0") }
        if constexpr (OP == 6) { IGC_PROBE_4("v_add_u32 %0, %1, %0") }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

template <int OP>
static int probe_one(igc_ctx *ctx, uint32_t blocks, float *ms_out)
{
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_valu_probe<OP>, dim3(blocks), dim3(256), 0, s, ctx->rcount.as<uint32_t>(), 7u);
    HIP_TRY(hipEventRecord(ctx->ev[0], s));
    for (int r = 0; r < 5; ++r)
        hipLaunchKernelGGL(k_valu_probe<OP>, dim3(blocks), dim3(256), 0, s, ctx->rcount.as<uint32_t>(), 8u + r);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx->ev[1], s));
    HIP_TRY(hipEventSynchronize(ctx->ev[1]));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    *ms_out = ms / 5;
    return IGC_OK;
}

} // namespace igc

using namespace igc;

extern "C" IGC_API int igc_valu_probe(igc_ctx *ctx, double *wave_instr_per_s /* [7] */)
{
    if (!ctx || !wave_instr_per_s) return set_error(IGC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t blocks = (uint32_t)ctx->cu_count * 32;
    int rc;
    if ((rc = ctx->rcount.ensure(sizeof(uint32_t) * (size_t)blocks * 256))) return rc;
    float ms[7];
    if ((rc = probe_one<0>(ctx, blocks, &ms[0])) || (rc = probe_one<1>(ctx, blocks, &ms[1])) ||
        (rc = probe_one<2>(ctx, blocks, &ms[2])) || (rc = probe_one<3>(ctx, blocks, &ms[3])) ||
        (rc = probe_one<4>(ctx, blocks, &ms[4])) || (rc = probe_one<5>(ctx, blocks, &ms[5])) ||
        (rc = probe_one<6>(ctx, blocks, &ms[6])))
        return rc;
    const double instr = (double)blocks * 4 * kProbeIters * 32; // wave-level VALU instructions per launch
    for (int k = 0; k < 7; ++k) wave_instr_per_s[k] = instr / (ms[k] * 1e-3);
    return IGC_OK;
}
