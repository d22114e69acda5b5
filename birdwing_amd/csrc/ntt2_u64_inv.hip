// ntt2_u64_inv.hip -- ArU64x inv register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_u64_inv(int logT, bool cols, int ingress) { return pick2<ArU64x, BW_MODE_INV>(logT, cols, ingress); }
