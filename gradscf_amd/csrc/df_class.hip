// Block kernels of the density-fitting integrals, one auxiliary angular
// momentum per translation unit (built with -DGSCF_DF_LC=0..4): orbital pairs
// (LA, LB) with LA <= 2, LB <= LA, and auxiliary "pairs" (LA, s0) with LA <= 4
// (an auxiliary shell times the unit function), against one auxiliary shell
// (LC, s0).
#include "jk_kernels.h"

#define GSCF_DF_CAT(a, b) a##b
#define GSCF_DF_NAME(p, c) GSCF_DF_CAT(p, c)

namespace gscf {

template <int LA, int LB>
static bool df_try(int la, int lb, const BlockArgs& a, hipStream_t s) {
  if (la != LA || lb != LB) return false;
  launch_blocks_t<LA, LB, GSCF_DF_LC, 0>(a, s);
  return true;
}

bool GSCF_DF_NAME(launch_df_blocks_c, GSCF_DF_LC)(int la, int lb, const BlockArgs& a, hipStream_t s) {
  return df_try<0, 0>(la, lb, a, s) || df_try<1, 0>(la, lb, a, s) || df_try<1, 1>(la, lb, a, s) ||
         df_try<2, 0>(la, lb, a, s) || df_try<2, 1>(la, lb, a, s) || df_try<2, 2>(la, lb, a, s) ||
         df_try<3, 0>(la, lb, a, s) || df_try<4, 0>(la, lb, a, s);
}

}  // namespace gscf
