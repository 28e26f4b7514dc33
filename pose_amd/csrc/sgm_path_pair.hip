// Multi-direction path launches (packed 16-bit sweeps reading a cost volume,
// lane groups 2, 4, 8): two directions per launch over two partial aggregate
// volumes (paired sweeps), and the middle directions of the path-cost plane
// schedule in one launch.
#include "sgm_internal.h"

namespace sgm {

template <int R, int PHB, int MODE>
static void launch_pair_t(hipStream_t st, const PathJobs &ja, const PathArgs &A,
                          const PathJobs &jb, const PathArgs &B, int waves) {
  const unsigned blocks = (unsigned)((A.total_lines + B.total_lines + waves - 1) / waves);
  path_pair_kernel<R, PHASE_MID, PHB, MODE><<<blocks, waves * 64, 0, st>>>(ja, A, jb, B);
}

template <int R, int PHB>
static int launch_pair_mode(hipStream_t st, int mode, const PathJobs &ja, const PathArgs &A,
                            const PathJobs &jb, const PathArgs &B, int waves) {
  if (mode == 4) launch_pair_t<R, PHB, 4>(st, ja, A, jb, B, waves);
  else if (mode == (4 | MODE_NT)) launch_pair_t<R, PHB, 4 | MODE_NT>(st, ja, A, jb, B, waves);
  else if (mode == 1) launch_pair_t<R, PHB, 1>(st, ja, A, jb, B, waves);
  else if (mode == (1 | MODE_NT)) launch_pair_t<R, PHB, 1 | MODE_NT>(st, ja, A, jb, B, waves);
  else return fail(SGM_ERR_INVALID_ARGUMENT, "no paired sweep for mode %d", mode);
  return launch_check("paired path kernel");
}

template <int R>
static int launch_pair_r(hipStream_t st, int mode, int phase_b, const PathJobs &ja,
                         const PathArgs &A, const PathJobs &jb, const PathArgs &B, int waves) {
  if (phase_b == PHASE_FIRST) return launch_pair_mode<R, PHASE_FIRST>(st, mode, ja, A, jb, B, waves);
  return launch_pair_mode<R, PHASE_MID>(st, mode, ja, A, jb, B, waves);
}

int launch_path_pair(sgm_context *ctx, int R, int mode, int phase_b, const PathJobs &ja,
                     const PathArgs &A, const PathJobs &jb, const PathArgs &B, int waves,
                     hipStream_t stream) {
  ProfileScope ps(ctx, "path_pair");
  if (A.total_lines + B.total_lines <= 0) return SGM_OK;
  hipStream_t st = stream ? stream : ctx->stream;
  switch (R) {
    case 2: return launch_pair_r<2>(st, mode, phase_b, ja, A, jb, B, waves);
    case 4: return launch_pair_r<4>(st, mode, phase_b, ja, A, jb, B, waves);
    case 8: return launch_pair_r<8>(st, mode, phase_b, ja, A, jb, B, waves);
    default: return fail(SGM_ERR_INVALID_ARGUMENT, "no paired sweep for lane group %d", R);
  }
}

template <int R, int MODE>
static void launch_multi_t(hipStream_t st, const PathJobs &jobs, const PathArgs &A,
                           const MultiDir &M, int waves) {
  const unsigned blocks = (unsigned)((M.total_lines + waves - 1) / waves);
  path_multi_kernel<R, MODE><<<blocks, waves * 64, 0, st>>>(jobs, A, M);
}

int launch_path_multi(sgm_context *ctx, int R, int mode, const PathJobs &jobs, const PathArgs &A,
                      const MultiDir &M, int waves, hipStream_t stream) {
  ProfileScope ps(ctx, "path_multi");
  if (M.total_lines <= 0) return SGM_OK;
  hipStream_t st = stream ? stream : ctx->stream;
  const bool nt = (mode & MODE_NT) != 0;
  if ((mode & ~MODE_NT) != (4 | MODE_PLANE))
    return fail(SGM_ERR_INVALID_ARGUMENT, "no multi-direction sweep for mode %d", mode);
  switch (R) {
    case 2:
      nt ? launch_multi_t<2, 4 | MODE_NT | MODE_PLANE>(st, jobs, A, M, waves)
         : launch_multi_t<2, 4 | MODE_PLANE>(st, jobs, A, M, waves);
      break;
    case 4:
      nt ? launch_multi_t<4, 4 | MODE_NT | MODE_PLANE>(st, jobs, A, M, waves)
         : launch_multi_t<4, 4 | MODE_PLANE>(st, jobs, A, M, waves);
      break;
    case 8:
      nt ? launch_multi_t<8, 4 | MODE_NT | MODE_PLANE>(st, jobs, A, M, waves)
         : launch_multi_t<8, 4 | MODE_PLANE>(st, jobs, A, M, waves);
      break;
    default:
      return fail(SGM_ERR_INVALID_ARGUMENT, "no multi-direction sweep for lane group %d", R);
  }
  return launch_check("multi-direction path kernel");
}

}  // namespace sgm
