// Status strings, version, and the synthetic-input generator.
//
// Generator: counter-based SplitMix64.  key = mix(seed); the i-th 64-bit
// word is mix(key + i), with mix(x) the SplitMix64 step (add the golden
// gamma, then the variant-13 finaliser).  u = (word >> 11) * 2^-53 in [0,1).
// The same definition is in tmv_amd/rng.py; uniform values are bit-identical
// (the affine map is computed without contraction).
#include "tmvx_common.h"
#include "tmvx_lu.h"
#include <mutex>
#include <vector>

namespace tmvx {

// Side streams of the look-ahead factorisation: a per-device free list of
// (stream, two events) sets, created on first need and kept until the
// process ends, so a call creates nothing once the pool holds as many sets
// as calls run at once.  Default stream priority.
namespace {
std::mutex g_side_mu;
std::vector<std::vector<SideSet*>> g_side_free;   // [device]
}

int side_acquire(SideSet** out)
{
    *out = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TMVX_ERR_HIP;
    {
        std::lock_guard<std::mutex> lk(g_side_mu);
        if ((int)g_side_free.size() <= dev) g_side_free.resize(dev + 1);
        auto& fl = g_side_free[dev];
        if (!fl.empty()) { *out = fl.back(); fl.pop_back(); return TMVX_OK; }
    }
    SideSet* ss = new SideSet;
    ss->device = dev;
    if (hipStreamCreateWithFlags(&ss->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ss->ev_main, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ss->ev_side, hipEventDisableTiming) != hipSuccess) {
        if (ss->ev_main) hipEventDestroy(ss->ev_main);
        if (ss->stream) hipStreamDestroy(ss->stream);
        delete ss;
        return TMVX_ERR_HIP;
    }
    *out = ss;
    return TMVX_OK;
}

void side_release(SideSet* ss)
{
    if (!ss) return;
    std::lock_guard<std::mutex> lk(g_side_mu);
    g_side_free[ss->device].push_back(ss);
}

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void fill_uniform_kernel(uint64_t key, uint64_t offset, i64 count,
                                    double lo, double span, double* out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t w = splitmix64(key + offset + (uint64_t)i);
    const double u = (double)(w >> 11) * 0x1.0p-53;
    out[i] = __dadd_rn(lo, __dmul_rn(span, u));
}

__global__ void fill_normal_kernel(uint64_t key, uint64_t offset, i64 count, double* out)
{
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t c = 2 * (offset + (uint64_t)i);
    const uint64_t w1 = splitmix64(key + c), w2 = splitmix64(key + c + 1);
    const double u1 = ((double)(w1 >> 11) + 0.5) * 0x1.0p-53;
    const double u2 = (double)(w2 >> 11) * 0x1.0p-53;
    out[i] = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

} // namespace tmvx

extern "C" {

const char* tmvx_status_str(int s)
{
    switch (s) {
    case TMVX_OK: return "ok";
    case TMVX_ERR_ARG: return "invalid argument";
    case TMVX_ERR_ALLOC: return "device allocation failed";
    case TMVX_ERR_HIP: return "HIP runtime or launch error";
    case TMVX_ERR_SINGULAR: return "singular triangular factor";
    case TMVX_ERR_NOT_CONVERGED: return "iterative refinement did not converge";
    case TMVX_ERR_TIMEOUT: return "bounded device wait expired";
    case TMVX_ERR_UNSUPPORTED: return "unsupported size or option";
    default: return "unknown status";
    }
}

const char* tmvx_version(void) { return "tmvx 0.1 (gfx950)"; }

int tmvx_fill_uniform_f64(uint64_t seed, uint64_t offset, int64_t count, double lo,
                          double hi, double* out, void* stream)
{
    if (count < 0 || (count > 0 && !out)) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    const uint64_t key = tmvx::splitmix64(seed);
    hipLaunchKernelGGL(tmvx::fill_uniform_kernel, dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, key, offset, count, lo, hi - lo,
                       out);
    return tmvx::hip_status(hipGetLastError());
}

int tmvx_fill_normal_f64(uint64_t seed, uint64_t offset, int64_t count, double* out,
                         void* stream)
{
    if (count < 0 || (count > 0 && !out)) return TMVX_ERR_ARG;
    if (count == 0) return TMVX_OK;
    const uint64_t key = tmvx::splitmix64(seed);
    hipLaunchKernelGGL(tmvx::fill_normal_kernel, dim3((unsigned)((count + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, key, offset, count, out);
    return tmvx::hip_status(hipGetLastError());
}

} // extern "C"
