// awsem_resident_f32.hip -- resident engine instantiated for float pair
// arithmetic (own translation unit so the two precisions compile in parallel).
#include "awsem_dev.h"

void aw_resident_prepare_f32()
{
  prepare_resident_t<float>();
}

int aw_launch_resident_f32(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  return launch_resident_t<float>(p, run, st);
}
