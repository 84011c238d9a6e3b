// FP32 instantiations of the model kernels for the mixed-precision path.  This file is
// compiled with approximate reciprocal and transcendental functions (division becomes
// v_rcp_f32 + multiply, exp/log become v_exp_f32/v_log_f32 with a scale), which removes
// about a third of the VALU instructions of the ORd step; the effect on Vm and on
// activation times is bounded by the mixed-vs-fp64 tissue tests.
#include "model_kernels.hpp"

namespace cg {
CG_DEFINE_MODEL_ENTRIES(float, fp32)
} // namespace cg

#include "batch_kernels.hpp"
namespace cg {
CG_DEFINE_BATCH_ENTRY(float, fp32)
} // namespace cg
