// ntt_f64_fwd.hip -- ArF64 fwd pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_f64_fwd(int logT, int G) { return pick_logT<ArF64, BW_MODE_FWD>(logT, G); }
