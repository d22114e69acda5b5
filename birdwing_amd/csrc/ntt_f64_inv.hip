// ntt_f64_inv.hip -- ArF64 inv pass kernels (instantiations of ntt_pass_impl.h).
#include "ntt_pass_impl.h"

PassKernel bw_pick_f64_inv(int logT, int G) { return pick_logT<ArF64, BW_MODE_INV>(logT, G); }
