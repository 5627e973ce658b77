// Tile variant 1 of the batched 2-D solver (kernels in batch2d.hpp).
#include "batch2d.hpp"

int lg::b2d::launch_v1(const hgrid_b2d_job *jobs, int njobs, int *queue, int nthreads, int per_cu, int nblocks) {
  if (nthreads == 512 && per_cu == 1) return launch<1, 512, 1>(jobs, njobs, queue, nblocks);
  if (nthreads == 512 && per_cu == 2) return launch<1, 512, 2>(jobs, njobs, queue, nblocks);
  if (nthreads == 1024 && per_cu == 1) return launch<1, 1024, 1>(jobs, njobs, queue, nblocks);
  return -1;
}
