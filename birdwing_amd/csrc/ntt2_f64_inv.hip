// ntt2_f64_inv.hip -- ArF64x inv register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_f64_inv(int logT, bool cols, int ingress) { return pick2<ArF64x, BW_MODE_INV>(logT, cols, ingress); }
