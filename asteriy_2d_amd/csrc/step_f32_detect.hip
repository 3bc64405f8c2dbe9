// step_f32_detect.hip -- the fast step kernels of the exact solver that flag
// strong two-shock faces, the other half of a gated learned step
// (step_f32_kernel.hpp).
#define AST_STEP_F32_KERNELS
#include "step_f32_kernel.hpp"

namespace ast {
namespace f32 {
template int step_launch<kExactDetect>(const StepArgs&);
}  // namespace f32
}  // namespace ast
