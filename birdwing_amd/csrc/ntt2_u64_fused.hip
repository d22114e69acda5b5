// ntt2_u64_fused.hip -- ArU64x fused register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_u64_fused(int logT, bool cols, int ingress) { return pick2<ArU64x, BW_MODE_FUSED>(logT, cols, ingress); }
