// ntt_u64_fwd.hip -- ArU64 fwd pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_u64_fwd(int logT, int G) { return pick_logT<ArU64, BW_MODE_FWD>(logT, G); }
