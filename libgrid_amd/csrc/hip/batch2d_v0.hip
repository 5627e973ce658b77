// Tile variant 0 of the batched 2-D solver (kernels in batch2d.hpp).
#include "batch2d.hpp"

int lg::b2d::launch_v0(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int nblocks) {
  if (nthreads == 256) return launch<0, 256, 1>(jobs, njobs, queue, nblocks);
  if (nthreads == 512) return launch<0, 512, 1>(jobs, njobs, queue, nblocks);
  return -1;
}
