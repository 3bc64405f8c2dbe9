// step_f32_learned.hip -- the fast step kernels of the exact solver with the
// learned start and of its strong-face detector (step_f32_kernel.hpp).
#define AST_STEP_F32_KERNELS
#include "step_f32_kernel.hpp"

namespace ast {
namespace f32 {
template int step_launch<kExactLearned>(const StepArgs&);
template int step_launch<kExactDetect>(const StepArgs&);
}  // namespace f32
}  // namespace ast
