// gpu_stub.cpp -- TEST INFRASTRUCTURE: the device entry points of klt.h as
// stubs that report "no GPU", so that the klitho_* tools build with the host
// code alone under AddressSanitizer / UndefinedBehaviorSanitizer
// (tests/test_host_sanitize.py) and run their argument and file handling up
// to the first device call.
#include "klt_error.h"

extern "C" {

int klt_ctx_create(int, klt_ctx** out) {
  if (out) *out = nullptr;
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
void klt_ctx_destroy(klt_ctx*) {}
int klt_tcc_build(klt_ctx*, const float*, int, float, float, float, float, float, int, int, int,
                  int, int, double*, unsigned) {
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
int klt_hopkins_image(klt_ctx*, const float*, int, int, const double*, int, int, float*,
                      unsigned) {
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
int klt_calc_kernels(klt_ctx*, const double*, int, int, float*, float*, float*, unsigned) {
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
int klt_kernels_create(klt_ctx*, const float*, const float*, const float*, int, int, int,
                       unsigned, klt_kernels** out) {
  if (out) *out = nullptr;
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
void klt_kernels_destroy(klt_kernels*) {}
int klt_socs_image(klt_ctx*, const klt_kernels*, const float*, int, int, float*, unsigned) {
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}
int klt_socs_forward(klt_ctx*, const klt_kernels*, const float*, int, int, int, float*,
                     unsigned) {
  return klt::set_error(KLT_ERR_NODEV, "no HIP device visible");
}

}  // extern "C"
