// One angular-momentum class of the J/K and block kernels per translation unit
// (built once per class with -DGSCF_LA/LB/LC/LD so the build parallelises).
#include "jk_kernels.h"
#include "grad_kernels.h"
#include "param_kernels.h"

#define GSCF_CAT5(a, b, c, d, e) a##b##c##d##e
#define GSCF_NAME(p, A, B, C, D) GSCF_CAT5(p, A, B, C, D)

namespace gscf {
void GSCF_NAME(launch_jk_c, GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD)(const JKArgs& a, hipStream_t s) {
  launch_jk_t<GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD>(a, s);
}
bool GSCF_NAME(launch_jk_bra_c, GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD)(const BraArgs& a, int pb, int pk, size_t lds,
                                                                   hipStream_t s) {
  return launch_jk_bra_patterns<GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD>(a, pb, pk, lds, s);
}
void GSCF_NAME(launch_grad_c, GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD)(const GradArgs& a, hipStream_t s) {
  launch_grad_t<GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD>(a, s);
}
void GSCF_NAME(launch_param_c, GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD)(const ParamArgs& a, hipStream_t s) {
  launch_param_t<GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD>(a, s);
}
void GSCF_NAME(launch_blocks_c, GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD)(const BlockArgs& a, hipStream_t s) {
  launch_blocks_t<GSCF_LA, GSCF_LB, GSCF_LC, GSCF_LD>(a, s);
}
}  // namespace gscf
