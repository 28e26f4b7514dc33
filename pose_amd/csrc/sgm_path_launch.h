// Path-sweep launch dispatch, shared by the three translation units that hold
// the sweep kernel's instances (first, middle and last sweep phases; each unit
// compiles one phase, so the instances build in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include "sgm_internal.h"

namespace sgm {

template <int R, typename SumT, int MODE, int PH>
static void launch_path_t(hipStream_t st, const PathJobs &jobs, const PathArgs &A, int waves) {
  const int threads = waves * 64;
  const unsigned blocks = (unsigned)((A.total_lines + waves - 1) / waves);
  path_kernel<R, SumT, PH, MODE><<<blocks, threads, 0, st>>>(jobs, A);
}

// MODE 0: generic 32-bit sweep (32-bit aggregates); 1: packed 16-bit; 2: packed
// 16-bit first sweep with the census cost fused in.
template <int PH>
static int launch_path_phase(hipStream_t st, int R, int sum_bytes, int mode, const PathJobs &jobs,
                             const PathArgs &A, int waves) {
#define CASE_WIDE(RR)                                                       \
  case RR:                                                                  \
    launch_path_t<RR, int32_t, 0, PH>(st, jobs, A, waves);              \
    break;
#define CASE_PK(RR)                                                         \
  case RR:                                                                  \
    if (mode == 2) launch_path_t<RR, uint16_t, 2, PH>(st, jobs, A, waves); \
    else if (mode == (2 | MODE_NT)) launch_path_t<RR, uint16_t, 2 | MODE_NT, PH>(st, jobs, A, waves); \
    else if (mode == 3) launch_path_t<RR, uint16_t, 3, PH>(st, jobs, A, waves); \
    else if (mode == 4) launch_path_t<RR, uint16_t, 4, PH>(st, jobs, A, waves); \
    else if (mode == (4 | MODE_NT)) launch_path_t<RR, uint16_t, 4 | MODE_NT, PH>(st, jobs, A, waves); \
    else if (mode == (1 | MODE_NT)) launch_path_t<RR, uint16_t, 1 | MODE_NT, PH>(st, jobs, A, waves); \
    else launch_path_t<RR, uint16_t, 1, PH>(st, jobs, A, waves);        \
    break;
  if constexpr (PH == PHASE_COST) {
    // the census cost pass (fused census cost, packed lane groups 2, 4, 8)
    if (sum_bytes != 2 || (mode & ~MODE_NT) != 2)
      return fail(SGM_ERR_INVALID_ARGUMENT, "the cost pass takes the fused census cost");
    const bool nt = (mode & MODE_NT) != 0;
    switch (R) {
      case 2: nt ? launch_path_t<2, uint16_t, 2 | MODE_NT, PH>(st, jobs, A, waves)
                 : launch_path_t<2, uint16_t, 2, PH>(st, jobs, A, waves); break;
      case 4: nt ? launch_path_t<4, uint16_t, 2 | MODE_NT, PH>(st, jobs, A, waves)
                 : launch_path_t<4, uint16_t, 2, PH>(st, jobs, A, waves); break;
      case 8: nt ? launch_path_t<8, uint16_t, 2 | MODE_NT, PH>(st, jobs, A, waves)
                 : launch_path_t<8, uint16_t, 2, PH>(st, jobs, A, waves); break;
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
    return launch_check("census cost pass");
  } else {
  if (mode & MODE_PLANE) {
    // path-cost planes (packed lane groups 2, 4, 8): first-phase sweeps store
    // their plane (fused census cost or uint8 cost volume), last sweeps add them
    const int m = mode & ~MODE_PLANE;
    if (sum_bytes != 2) return fail(SGM_ERR_INVALID_ARGUMENT, "path-cost planes need 16-bit sums");
    if (R != 2 && R != 4 && R != 8) return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    if constexpr (PH == PHASE_MID) {
      return fail(SGM_ERR_INVALID_ARGUMENT, "no middle sweep over path-cost planes");
    } else {
      constexpr int P = MODE_PLANE;
#define PLANE_R(MM)                                                                   \
  (R == 2 ? (launch_path_t<2, uint16_t, MM, PH>(st, jobs, A, waves), 0)               \
   : R == 4 ? (launch_path_t<4, uint16_t, MM, PH>(st, jobs, A, waves), 0)             \
            : (launch_path_t<8, uint16_t, MM, PH>(st, jobs, A, waves), 0))
      if (m == 4) (void)PLANE_R(4 | P);
      else if (m == (4 | MODE_NT)) (void)PLANE_R(4 | MODE_NT | P);
      else if constexpr (PH == PHASE_FIRST) {
        if (m == 2) (void)PLANE_R(2 | P);
        else if (m == (2 | MODE_NT)) (void)PLANE_R(2 | MODE_NT | P);
        else return fail(SGM_ERR_INVALID_ARGUMENT, "no path-cost plane sweep for mode %d", m);
      } else {
        return fail(SGM_ERR_INVALID_ARGUMENT, "no path-cost plane sweep for mode %d", m);
      }
#undef PLANE_R
      return launch_check("path kernel");
    }
  }
  if (mode & MODE_SUM2) {
    // last sweep over paired partial aggregates (packed lane groups 2, 4, 8)
    if constexpr (PH == PHASE_LAST) {
      const int m = mode & ~MODE_SUM2;
#define CASE_SUM2(RR)                                                                     \
  case RR:                                                                                \
    if (m == 1) launch_path_t<RR, uint16_t, 1 | MODE_SUM2, PH>(st, jobs, A, waves);       \
    else if (m == (1 | MODE_NT)) launch_path_t<RR, uint16_t, 1 | MODE_NT | MODE_SUM2, PH>(st, jobs, A, waves); \
    else if (m == 4) launch_path_t<RR, uint16_t, 4 | MODE_SUM2, PH>(st, jobs, A, waves);  \
    else if (m == (4 | MODE_NT)) launch_path_t<RR, uint16_t, 4 | MODE_NT | MODE_SUM2, PH>(st, jobs, A, waves); \
    else return fail(SGM_ERR_INVALID_ARGUMENT, "no paired-aggregate sweep for mode %d", m); \
    break;
      if (sum_bytes != 2) return fail(SGM_ERR_INVALID_ARGUMENT, "paired aggregates are 16-bit");
      switch (R) {
        CASE_SUM2(2) CASE_SUM2(4) CASE_SUM2(8)
        default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
      }
#undef CASE_SUM2
      return launch_check("path kernel");
    } else {
      return fail(SGM_ERR_INVALID_ARGUMENT, "a second partial aggregate is read by the last sweep");
    }
  }
  if (sum_bytes == 4) {
    switch (R) {
      CASE_WIDE(1) CASE_WIDE(2) CASE_WIDE(4) CASE_WIDE(8) CASE_WIDE(16) CASE_WIDE(32) CASE_WIDE(64)
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  } else {
    switch (R) {
      CASE_PK(2) CASE_PK(4) CASE_PK(8)
#define CASE_PK_LOAD(RR)                                                    \
  case RR:                                                                  \
    if ((mode & 7) == 4) launch_path_t<RR, uint16_t, 4, PH>(st, jobs, A, waves); \
    else launch_path_t<RR, uint16_t, 1, PH>(st, jobs, A, waves);        \
    break;
      CASE_PK_LOAD(16) CASE_PK_LOAD(32) CASE_PK_LOAD(64)
#undef CASE_PK_LOAD
      default: return fail(SGM_ERR_INVALID_ARGUMENT, "unsupported lane group %d", R);
    }
  }
#undef CASE_WIDE
#undef CASE_PK
  return launch_check("path kernel");
  }
}

}  // namespace sgm
