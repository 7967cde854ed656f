// awsem_resident_f64.hip -- resident engine instantiated for double pair
// arithmetic (own translation unit so the two precisions compile in parallel).
#include "awsem_dev.h"

void aw_resident_prepare_f64()
{
  prepare_resident_t<double>();
}

int aw_launch_resident_f64(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  return launch_resident_t<double>(p, run, st);
}
