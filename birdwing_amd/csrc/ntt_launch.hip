// ntt_launch.hip -- kernel selection and launch of one NTT pass.
#include "ntt_kernels.h"

typedef void (*PassKernel)(PassParams);
PassKernel bw_pick_f64_fwd(int logT, int G);
PassKernel bw_pick_f64_inv(int logT, int G);
PassKernel bw_pick_f64_fused(int logT, int G);
PassKernel bw_pick_u64_fwd(int logT, int G);
PassKernel bw_pick_u64_inv(int logT, int G);
PassKernel bw_pick_u64_fused(int logT, int G);

int bw_pass_max_G(int logT) {
    int g = 16;
    while (g > 1 && (1 << logT) * g > 8192) g >>= 1;
    return g;
}

hipError_t bw_launch_pass(const PassLaunch& L, const PassParams& prm, hipStream_t stream) {
    PassKernel k = nullptr;
    if (L.limb_bits == 64) {
        if (L.mode == BW_MODE_FWD) k = bw_pick_u64_fwd(L.logT, L.G);
        else if (L.mode == BW_MODE_INV) k = bw_pick_u64_inv(L.logT, L.G);
        else k = bw_pick_u64_fused(L.logT, L.G);
    } else {
        if (L.mode == BW_MODE_FWD) k = bw_pick_f64_fwd(L.logT, L.G);
        else if (L.mode == BW_MODE_INV) k = bw_pick_f64_inv(L.logT, L.G);
        else k = bw_pick_f64_fused(L.logT, L.G);
    }
    if (!k) return hipErrorInvalidValue;
    const size_t vbytes = 8;   // FP64 and u64 register values are both 8 bytes
    const size_t lds = (size_t)L.G * lds_line(L.logT) * vbytes;
    static bool attr_set[2][3][16][17] = {};
    bool& done = attr_set[L.limb_bits == 64][L.mode][L.logT][L.G];
    if (!done) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        done = true;
    }
    dim3 grid((unsigned)L.tiles, (unsigned)L.nprimes, (unsigned)(L.mode == BW_MODE_FWD ? L.nops : 1));
    hipLaunchKernelGGL(k, grid, dim3(256), lds, stream, prm);
    return hipGetLastError();
}
