// ntt2_u64_fwd.hip -- ArU64x fwd register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_u64_fwd(int logT, bool cols, int ingress) { return pick2<ArU64x, BW_MODE_FWD>(logT, cols, ingress); }
