// Three-body kernels (Stillinger-Weber, Tersoff).
#include <hip/hip_runtime.h>

#include "omdg_internal.h"

namespace omdg {

void sw_forces(Ctx&, bool) {
  throw ApiFail(OMDG_EINVAL, "Stillinger-Weber kernel not built yet");
}

void tersoff_forces(Ctx&, bool) {
  throw ApiFail(OMDG_EINVAL, "Tersoff kernel not built yet");
}

}  // namespace omdg
