// step_sources.h -- the neutral-H source planes of one FP32 step, as the
// runtime and the step's C entry point hand them to the step's launcher.
#pragma once

namespace ast {
struct StepSources {
  const double* nn2;  // the reference's FP64 planes: N*M triples (x, y used)
  const double* nn3;  // and N*M values; either may be null
  const float* xy;    // or FP32 copies, read instead when either is set:
  const float* e;     // nn2's (x, y) as N*M pairs, nn3 as N*M values
  double kmom, ken;
};
}  // namespace ast
