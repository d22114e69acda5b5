// ntt2_f64_fwd.hip -- ArF64x fwd register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_f64_fwd(int logT, bool cols, int ingress) { return pick2<ArF64x, BW_MODE_FWD>(logT, cols, ingress); }
