// Instances of the path-sweep kernel for the first sweep phase.
#include "sgm_path_launch.h"

namespace sgm {

int launch_path_first(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                      const PathArgs &A, int waves) {
  return launch_path_phase<PHASE_FIRST>(st, R, sum_bytes, mode, jobs, A, waves);
}

int launch_path_cost(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                     const PathArgs &A, int waves) {
  return launch_path_phase<PHASE_COST>(st, R, sum_bytes, mode, jobs, A, waves);
}

}  // namespace sgm
