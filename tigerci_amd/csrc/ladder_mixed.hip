// Split-precision ladder (BF16x3 / scaled FP16x3) -- filled in below.
#include "plan.hpp"

namespace tci {

void mixed_prepare_L(tci_plan& P, hipStream_t st) {
  (void)P;
  (void)st;
  throw Error(TCI_ERR_ARG, "mixed-precision ladder not built yet");
}

void mixed_ladder(tci_plan& P, int precision, hipStream_t st) {
  (void)P;
  (void)precision;
  (void)st;
  throw Error(TCI_ERR_ARG, "mixed-precision ladder not built yet");
}

}  // namespace tci
