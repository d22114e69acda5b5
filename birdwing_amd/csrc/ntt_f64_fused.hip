// ntt_f64_fused.hip -- ArF64 fused pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_f64_fused(int logT, int G) { return pick_logT<ArF64, BW_MODE_FUSED>(logT, G); }
