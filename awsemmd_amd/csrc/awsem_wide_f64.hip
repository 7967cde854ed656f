// awsem_wide_f64.hip -- wide engine instantiated for double pair arithmetic.
#include "awsem_wide.h"

int aw_launch_wide_f64(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st)
{
  return launch_wide_t<double>(p, run, max_steps, st);
}

void aw_wide_prepare_f64()
{
  prepare_wide_t<double>();
}
