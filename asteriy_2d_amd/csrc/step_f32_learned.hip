// step_f32_learned.hip -- the fast step kernels of the exact solver with the
// learned start (step_f32_kernel.hpp).
#define AST_STEP_F32_KERNELS
#include "step_f32_kernel.hpp"

namespace ast {
namespace f32 {
template int step_launch<kExactLearned>(const StepArgs&);
}  // namespace f32
}  // namespace ast
