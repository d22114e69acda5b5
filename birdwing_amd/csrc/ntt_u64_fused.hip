// ntt_u64_fused.hip -- ArU64 fused pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_u64_fused(int logT, int G) { return pick_logT<ArU64, BW_MODE_FUSED>(logT, G); }
