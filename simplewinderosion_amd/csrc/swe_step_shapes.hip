// swe_step_shapes.hip -- one tile shape of the step kernel per compilation:
// built once per shape with -DSWE_STEP_SHAPE=<index into kStepShape>, so
// the shapes compile in parallel.
#include "swe_step_kernel.hpp"

#ifndef SWE_STEP_SHAPE
#error "compile with -DSWE_STEP_SHAPE=<index into kStepShape>"
#endif

namespace swe {

#define SWE_CAT2(a, b) a##b
#define SWE_CAT(a, b) SWE_CAT2(a, b)
void SWE_CAT(launch_step_shape, SWE_STEP_SHAPE)(const StepArgs& a, bool periodic, bool track, int layers, bool deposit,
                                                int mode, hipStream_t s) {
  constexpr StepShape sh = kStepShape[SWE_STEP_SHAPE];
  StepLauncher<sh.sx, sh.sy, sh.threads, sh.waves>::launch(a, periodic, track, layers, deposit, mode, s);
}

}  // namespace swe
