// ntt2_f64_fused.hip -- ArF64x fused register-staged pass kernels (instantiations of ntt2_impl.h).
#include "ntt2_impl.h"

PassKernel2 bw_pick2_f64_fused(int logT, bool cols, int ingress) { return pick2<ArF64x, BW_MODE_FUSED>(logT, cols, ingress); }
