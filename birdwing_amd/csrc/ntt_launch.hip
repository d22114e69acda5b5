// ntt_launch.hip -- kernel selection and launch of one NTT pass.
#include "ntt_kernels.h"

#include <mutex>
#include <unordered_set>

typedef void (*PassKernel)(PassParams);
PassKernel bw_pick_f64_fwd(int logT, int G);
PassKernel bw_pick_f64_inv(int logT, int G);
PassKernel bw_pick_f64_fused(int logT, int G);
PassKernel bw_pick_u64_fwd(int logT, int G);
PassKernel bw_pick_u64_inv(int logT, int G);
PassKernel bw_pick_u64_fused(int logT, int G);
PassKernel bw_pick2_f64_fwd(int logT, bool cols, int ingress);
PassKernel bw_pick2_f64_inv(int logT, bool cols, int ingress);
PassKernel bw_pick2_f64_fused(int logT, bool cols, int ingress);
PassKernel bw_pick2_u64_fwd(int logT, bool cols, int ingress);
PassKernel bw_pick2_u64_inv(int logT, bool cols, int ingress);
PassKernel bw_pick2_u64_fused(int logT, bool cols, int ingress);

int bw_pass_max_G(int logT, int limb_bits, bool cols) {
    if (logT <= BW_REG2_MAX_LOG) {
        // register-staged: G * R1 <= 512 threads and at most 64 KiB of LDS
        // values (4-byte values on the 32-bit path, 8-byte on the 64-bit path)
        const int r1 = t2_r1log(logT, cols);
        const int cap = limb_bits == 64 ? 8192 : 16384;
        int g = 512 >> r1;
        while (g > 1 && (1 << logT) * g > cap) g >>= 1;
        return g;
    }
    int g = 16;
    while (g > 1 && (1 << logT) * g > 8192) g >>= 1;
    return g;
}

int bw_reg2_lds_line(int logT, bool cols) { return t2_lds_line(logT, cols); }

hipError_t bw_launch_pass(const PassLaunch& L, const PassParams& prm, hipStream_t stream) {
    if (L.mode == BW_MODE_FUSED && L.limb_bits == 32 && L.logT >= BW_ROW_MIN_LOG) {
        if (prm.ingress || prm.bdesc || L.logT > BW_ROW_MAX_LOG) return hipErrorInvalidValue;
        return bw_launch_row_fused(L.logT, L.tiles, L.nprimes, prm, stream);
    }
    PassKernel k = nullptr;
    const bool reg2 = L.logT <= BW_REG2_MAX_LOG;   // register-staged kernels
    if (reg2) {
        if (L.limb_bits == 64) {
            if (L.mode == BW_MODE_FWD) k = bw_pick2_u64_fwd(L.logT, prm.S > 1, prm.ingress);
            else if (L.mode == BW_MODE_INV) k = bw_pick2_u64_inv(L.logT, prm.S > 1, prm.ingress);
            else k = bw_pick2_u64_fused(L.logT, prm.S > 1, prm.ingress);
        } else {
            if (L.mode == BW_MODE_FWD) k = bw_pick2_f64_fwd(L.logT, prm.S > 1, prm.ingress);
            else if (L.mode == BW_MODE_INV) k = bw_pick2_f64_inv(L.logT, prm.S > 1, prm.ingress);
            else k = bw_pick2_f64_fused(L.logT, prm.S > 1, prm.ingress);
        }
    } else if (L.limb_bits == 64) {
        if (L.mode == BW_MODE_FWD) k = bw_pick_u64_fwd(L.logT, L.G);
        else if (L.mode == BW_MODE_INV) k = bw_pick_u64_inv(L.logT, L.G);
        else k = bw_pick_u64_fused(L.logT, L.G);
    } else {
        if (L.mode == BW_MODE_FWD) k = bw_pick_f64_fwd(L.logT, L.G);
        else if (L.mode == BW_MODE_INV) k = bw_pick_f64_inv(L.logT, L.G);
        else k = bw_pick_f64_fused(L.logT, L.G);
    }
    if (!k) return hipErrorInvalidValue;
    // LDS element: 8 bytes, except the register-staged 32-bit path (centered int32)
    const size_t vbytes = (reg2 && L.limb_bits == 32) ? 4 : 8;
    // long-stride inverse column passes: twice the tile width (the kernel's
    // tile and prime mapping follow blockDim)
    const bool wide_ok = reg2 && L.mode == BW_MODE_INV && prm.S > 1 && t2_inv_wide_ok(L.logT, (int)vbytes);
    PassLaunch LW = L;
    if (wide_ok && L.G == (512 >> 4) && prm.S >= ((int64_t)1 << BW_INV_WIDE_MIN_LOGS) && prm.S % (2 * L.G) == 0 &&
        L.tiles % 2 == 0) {
        LW.G = 2 * L.G;
        LW.tiles = L.tiles / 2;
    }
    const size_t lds = reg2 ? (size_t)t2_lds_bytes(L.logT, prm.S > 1, LW.G, (int)vbytes, L.mode, prm.ingress)
                            : (size_t)L.G * lds_line(L.logT) * vbytes;
    if (L.G > bw_pass_max_G(L.logT, L.limb_bits, prm.S > 1) || (L.G & (L.G - 1))) return hipErrorInvalidValue;
    if (prm.bdesc && prm.ingress && !reg2) return hipErrorInvalidValue;   // stacked ingress: reg2 only
    if (prm.ingress >= 2 && !reg2) return hipErrorInvalidValue;           // limb pairs / 32-bit limbs in 64-bit words
    // register-staged column tiles of 32-bit words address their T*S block
    // with 32-bit byte offsets from the tile's first element (8 bytes per
    // element for the raw 64-bit digits of the ingress pass)
    if (reg2 && L.limb_bits == 32 && prm.S > 1 && (prm.S << L.logT) > ((int64_t)1 << 29)) return hipErrorInvalidValue;
    {
        // Allow the largest tile this kernel is ever launched with, once per kernel.
        static std::mutex mu;
        static std::unordered_set<const void*> done;
        std::lock_guard<std::mutex> lock(mu);
        if (!done.count(reinterpret_cast<const void*>(k))) {
            const int gmax = bw_pass_max_G(L.logT, L.limb_bits, prm.S > 1) * (wide_ok ? 2 : 1);
            const size_t lds_max = reg2 ? (size_t)t2_lds_bytes(L.logT, prm.S > 1, gmax, (int)vbytes, L.mode, prm.ingress)
                                        : (size_t)gmax * lds_line(L.logT) * vbytes;
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max);
            if (e != hipSuccess) return e;
            done.insert(reinterpret_cast<const void*>(k));
        }
    }
    dim3 grid((unsigned)LW.tiles, (unsigned)L.nprimes, (unsigned)(L.mode == BW_MODE_FWD ? L.nops : 1));
    // register-staged kernels run G * R1 threads (one stage-0 group each)
    const unsigned nt = reg2 ? (unsigned)(LW.G << t2_r1log(L.logT, prm.S > 1)) : 256u;
    if (nt > (wide_ok ? 1024u : 512u) || nt == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, grid, dim3(nt), lds, stream, prm);
    return hipGetLastError();
}
