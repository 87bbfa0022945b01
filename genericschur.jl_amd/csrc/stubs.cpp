// Entry points declared in the ABI whose kernels are not built yet.
#include "../../include/gschur_mi355x.h"

extern "C" {
int gs_trevc_z(gs_handle*, int64_t, const double*, int64_t, const double*, int64_t, double*, int64_t, int, int, int) {
  return GS_E_UNSUPPORTED;
}
int gs_trevc_d(gs_handle*, int64_t, const double*, int64_t, const double*, int64_t, double*, int64_t, int, int, int) {
  return GS_E_UNSUPPORTED;
}
int gs_trevc_batched_z(gs_handle*, int64_t, const int32_t*, const int64_t*, const double*, const double*, double*, int,
                       int, int) {
  return GS_E_UNSUPPORTED;
}
int gs_trevc_batched_d(gs_handle*, int64_t, const int32_t*, const int64_t*, const double*, const double*, double*, int,
                       int, int) {
  return GS_E_UNSUPPORTED;
}
int gs_eig_vjp_batched_d(gs_handle*, int64_t, const int32_t*, const int64_t*, const double*, const double*,
                         const double*, double*, int32_t*, int) {
  return GS_E_UNSUPPORTED;
}
}
