// ntt_u64_inv.hip -- ArU64 inv pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_u64_inv(int logT, int G) { return pick_logT<ArU64, BW_MODE_INV>(logT, G); }
