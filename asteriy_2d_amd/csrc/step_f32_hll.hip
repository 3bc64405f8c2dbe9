// step_f32_hll.hip -- the fast step kernels of the HLL, HLLC and HLLC
// (Aleksashov) fluxes (step_f32_kernel.hpp).
#define AST_STEP_F32_KERNELS
#include "step_f32_kernel.hpp"

namespace ast {
namespace f32 {
template int step_launch<AST_HLL>(const StepArgs&);
template int step_launch<AST_HLLC>(const StepArgs&);
template int step_launch<AST_HLLC_A>(const StepArgs&);
}  // namespace f32
}  // namespace ast
