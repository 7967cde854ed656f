// awsem_wide_f32.hip -- wide engine instantiated for float pair arithmetic.
#include "awsem_wide.h"

int aw_launch_wide_f32(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st)
{
  return launch_wide_t<float>(p, run, max_steps, st);
}

void aw_wide_prepare_f32()
{
  prepare_wide_t<float>();
}
