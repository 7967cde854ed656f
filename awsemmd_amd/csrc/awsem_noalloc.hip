// awsem_noalloc.hip -- awsem_compute_res_noalloc: the residue-layout force
// evaluation of awsem_compute_res with every per-call device buffer in one
// fixed, file-scope device array, so that the call allocates nothing.
//
// The array (`scratch`, 4 MiB on each GPU the library runs on, held until
// the process ends) holds the coordinate staging, the force output, the
// energies and the phase engine's working slab of one evaluation.  Host code
// never passes the array's name to a kernel or a copy: on each device, on the
// entry point's first call there, the array's device address is read with
// hipGetSymbolAddress (with that device current) and kept per device; every
// launch and copy uses that address.  Every call on one GPU shares its array,
// which keeps the last call's contents; the phase kernels overwrite or clear
// everything they read (coordinates copied in, forces and energies zeroed by
// the first phase), and calls that overlap on one GPU (other threads, other
// contexts) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/awsem.h"
#include "awsem_launch.h"

#define AW_NOALLOC_DOUBLES ((size_t)(4u << 20) / sizeof(double))
#define AW_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[AW_NOALLOC_DOUBLES];

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[AW_NOALLOC_MAX_DEVICES];

// the device address of `scratch` on the current device (looked up once per
// device); null on failure
static double *noalloc_region(int *rc)
{
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= AW_NOALLOC_MAX_DEVICES) {
    *rc = AWSEM_E_DEVICE;
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *rc = AWSEM_E_DEVICE;
      return nullptr;
    }
    g_noalloc_ptr[dev] = p;
  }
  *rc = AWSEM_OK;
  return (double *)g_noalloc_ptr[dev];
}

extern "C" int awsem_compute_res_noalloc(awsem_ctx *c, const double *xres, double *fres, double *energy)
{
  int rc = aw_noalloc_check(c, xres, fres, AW_NOALLOC_DOUBLES);
  if (rc != AWSEM_OK) return rc;
  double *region = noalloc_region(&rc);
  if (!region) return rc;
  return aw_compute_res_scratch(c, xres, fres, energy, region);
}
