// gpp_factored.hip -- factored form of the GPP sigma band loop (placeholder
// until the MFMA Gram-product kernels land; reports the form unavailable).
#include "gpp_internal.hpp"

namespace gpp {

size_t factored_workspace_bytes(const DevProblem &, bool) { return 0; }

hipError_t run_factored(const DevProblem &, bool, void *, const DevResults &, hipStream_t, StageTimer *)
{
    return hipErrorNotSupported;
}

}  // namespace gpp
