// sampler_batched.hip -- nnz-bucketed sampler (under construction).
#include <stdexcept>
#include "common.h"
#include "kernels.h"
namespace mgpu {
int batched_supported(int, int) { return 0; }
void launch_sample_batched(const SampleArgs&, const BatchPlan&, hipStream_t) {
  throw std::runtime_error("batched sampler not built");
}
}  // namespace mgpu
