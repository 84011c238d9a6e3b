// FP64 instantiations of the model kernels (IEEE division, exp and log): the plain
// double-precision path of the tissue step and of the single-cell runs.
#include "model_kernels.hpp"

namespace cg {
CG_DEFINE_MODEL_ENTRIES(double, fp64)
} // namespace cg

#include "batch_kernels.hpp"
namespace cg {
CG_DEFINE_BATCH_ENTRY(double, fp64)

const char* batch_variant_name(int v)
{
    static const char* names[BATCH_VARIANTS] = {"256 threads", "64 threads",
                                                "256 threads, >=2 waves/SIMD",
                                                "64 threads, >=2 waves/SIMD",
                                                "256 threads, >=3 waves/SIMD", "128 threads"};
    return v >= 0 && v < BATCH_VARIANTS ? names[v] : "?";
}
} // namespace cg
