// HBM stream probe: the copy path of the streaming contraction kernels without their arithmetic, to
// measure the rate this part sustains for a given read/write mix. Each block of read_kb KB is copied
// HBM -> LDS with global_load_lds_dwordx4 (non-temporal) through a ring of four 32 KB units (three in
// flight, one barrier per unit, one 512-thread workgroup per CU, as the contraction kernels run),
// and write_kb KB are stored per block with 16-byte lanes (the contraction writes its results once
// per block the same way).
#include <algorithm>

#include "xtr_common.h"
#include "xtr_kernels.h"

namespace xtr {
namespace {

constexpr int kProbeWaves = 8, kProbeUnitKb = 32, kProbeSlots = 4;
constexpr int kProbePerWave = kProbeUnitKb / kProbeWaves;   // wave-loads (1 KB) per wave per unit

template <int N>
__device__ __forceinline__ void probe_wait_vm()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__global__ __launch_bounds__(64 * kProbeWaves) void probe_stream_kernel(const double2 *__restrict__ src, double2 *dst,
                                                                        int64_t blocks, int read_kb, int write_kb)
{
    __shared__ double2 ring[kProbeSlots * kProbeUnitKb * 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) double2 *)ring;
    const int64_t g = gridDim.x;
    const int64_t cnt = blocks > blockIdx.x ? (blocks - 1 - blockIdx.x) / g + 1 : 0;
    const int nu = read_kb / kProbeUnitKb;                      // units per block
    const int64_t total = cnt * nu;
    auto issue = [&](int64_t u) {
        const int64_t b = blockIdx.x + (u / nu) * g;
        const double2 *p0 = src + (b * read_kb + (u % nu) * kProbeUnitKb) * 64;
        const unsigned d0 = base + (unsigned)((u % kProbeSlots) * kProbeUnitKb * 1024);
#pragma unroll
        for (int t = 0; t < kProbePerWave; ++t) {
            const int wl = wave + kProbeWaves * t;
            const double2 *p = p0 + wl * 64 + lane;
            const unsigned m0 = __builtin_amdgcn_readfirstlane(d0 + (unsigned)wl * 1024u);
            asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off nt" ::"s"(m0), "v"(p) : "memory", "m0");
        }
    };
    for (int64_t u = 0; u < kProbeSlots - 1 && u < total; ++u) issue(u);
    double acc = 0;
    for (int64_t u = 0; u < total; ++u) {
        if (u + kProbeSlots - 2 < total) probe_wait_vm<(kProbeSlots - 2) * kProbePerWave>();
        else probe_wait_vm<0>();
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        if (u + kProbeSlots - 1 < total) issue(u + kProbeSlots - 1);
        acc += ring[(u % kProbeSlots) * kProbeUnitKb * 64 + threadIdx.x].x;
        if (u % nu == nu - 1 && write_kb > 0) {
            const int64_t b = blockIdx.x + (u / nu) * g;
            double2 *q = dst + b * write_kb * 64;
            for (int e = threadIdx.x; e < write_kb * 64; e += 64 * kProbeWaves) q[e] = make_double2(acc, (double)e);
        }
    }
    probe_wait_vm<0>();                                         // no copy may land after exit
}
#pragma clang diagnostic pop

}  // namespace

void launch_probe_stream(const void *src, void *dst, int64_t blocks, int read_kb, int write_kb, hipStream_t s)
{
    if (blocks < 1 || read_kb < kProbeUnitKb || read_kb % kProbeUnitKb != 0 || write_kb < 0)
        throw hip_error("probe: blocks >= 1, read_kb a positive multiple of 32 and write_kb >= 0");
    int dev = 0, cus = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    const unsigned grid = (unsigned)std::min<int64_t>(blocks, cus);
    hipLaunchKernelGGL(probe_stream_kernel, dim3(grid), dim3(64 * kProbeWaves), 0, s, (const double2 *)src,
                       (double2 *)dst, blocks, read_kb, write_kb);
    XTR_LAUNCH_CHECK("probe_stream_kernel");
}

}  // namespace xtr
