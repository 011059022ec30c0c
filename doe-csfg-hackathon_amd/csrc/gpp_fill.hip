// gpp_fill.hip -- device generator for the seeded GPP inputs (bit-identical
// to the host generator gpp_fill_seeded: both evaluate gpp_rng.hpp).
#include "gpp_internal.hpp"
#include "gpp_rng.hpp"

namespace gpp {
namespace {

__global__ void k_fill_seeded(cplx *dst, size_t n, uint64_t seed, uint32_t array_id, double rlo, double rhi,
                              double ilo, double ihi)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        cplx v;
        v.re = seeded_uniform(seed, array_id, 2 * i, rlo, rhi);
        v.im = seeded_uniform(seed, array_id, 2 * i + 1, ilo, ihi);
        dst[i] = v;
    }
}

}  // namespace

hipError_t launch_fill_seeded(cplx *dst, size_t n, uint64_t seed, uint32_t array_id, double re_lo, double re_hi,
                              double im_lo, double im_hi, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_fill_seeded, dim3((unsigned)blocks), dim3(256), 0, s, dst, n, seed, array_id, re_lo, re_hi,
                       im_lo, im_hi);
    return hipGetLastError();
}

}  // namespace gpp
