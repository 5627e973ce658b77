// C-ABI of the batched 2-D imaginary-time solver (kernels in batch2d.hpp).
#include "batch2d.hpp"

extern "C" int hgrid_b2d_supported(long n) { return n == 64 || n == 96 || n == 128 || n == 192 || n == 256; }

extern "C" int hgrid_b2d_cus(void) {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  return n;
}

extern "C" int hgrid_b2d_launch(const hgrid_b2d_job *jobs, int njobs, int *queue, int variant, int nthreads, int per_cu,
                                int nblocks) {
  if (njobs <= 0) return 0;
  if (nblocks <= 0) return -1;
  if (lg::check(hipGetLastError(), "before batch2d")) return -1;
  if (hgrid_memset(queue, 0, sizeof(int))) return -1;
  if (variant == 0) return lg::b2d::launch_v0(jobs, njobs, queue, nthreads, per_cu, nblocks);
  if (variant == 1) return lg::b2d::launch_v1(jobs, njobs, queue, nthreads, per_cu, nblocks);
  return -1;
}
