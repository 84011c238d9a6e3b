// FP64 instantiations of the model kernels (IEEE division, exp and log): the plain
// double-precision path of the tissue step and of the single-cell runs.
#include "model_kernels.hpp"

namespace cg {
CG_DEFINE_MODEL_ENTRIES(double, fp64)
} // namespace cg
