// step_f32_exact.hip -- the fast step kernels of the exact solver (step_f32_kernel.hpp).
#define AST_STEP_F32_KERNELS
#include "step_f32_kernel.hpp"

namespace ast {
namespace f32 {
template int step_launch<AST_EXACT>(const StepArgs&);
}  // namespace f32
}  // namespace ast
