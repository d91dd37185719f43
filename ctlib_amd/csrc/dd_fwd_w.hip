// Stand-alone launch of the weighted line -> bin kernel with wave-private line
// tables (fwd_block_w in dd_device.h): backprojection_t and the exact adjoint of
// backprojection, for lines of at most 1024 pixels.
#include <hip/hip_runtime.h>

#include "dd_device.h"
#include "kernels.h"

namespace ctlib {

using namespace dd;

namespace {

// Work items [0, B*ngx) are (b, group of 4 X-axis views) reading image rows from
// `srcx`; the rest are Y-axis groups reading the bottom-to-top column copy.
// fp32 variants for 512-pixel lines are held to 128 VGPRs: four waves per
// SIMD, as many as the LDS of four 512-pixel tables per block allows.
template <typename T, int NEM, int PPL, bool FULL>
constexpr int fwd_w_waves() { return sizeof(T) == 4 && FULL && PPL == 8 && NEM <= 12 ? 4 : 1; }

template <typename T, int NEM, int PPL, bool FULL>
__global__ __launch_bounds__(256, (fwd_w_waves<T, NEM, PPL, FULL>())) void k_fwd_w(DevPlan P, const T* __restrict__ srcx, const T* __restrict__ srcy,
                                               T* __restrict__ y, int B, int ngx, int ngy,
                                               const double* __restrict__ post, int nch) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nwx = B * ngx;
    int w = xcd_remap(blockIdx.x, gridDim.x, (nwx + B * ngy) * nch);
    if (w < 0) return;
    const int ch = w % nch;
    w /= nch;
    const int mode = w < nwx ? MODE_X : MODE_Y;
    if (mode == MODE_Y) w -= nwx;
    const int ngroups = mode == MODE_Y ? ngy : ngx;
    const int b = w / ngroups, grp = w - b * ngroups;
    const int nl = mode == MODE_Y ? P.W : P.H;
    fwd_block_w<T, NEM, PPL, FULL>(P, mode == MODE_Y ? srcy : srcx, y + (size_t)ch * B * P.V * P.D, b, mode, grp,
                                   post, smem, ch * nl / nch, (ch + 1) * nl / nch);
}

int line_max(const DevPlan& P) {
    int n = P.nx ? P.W : 0;
    if (P.ny && P.H > n) n = P.H;
    return n;
}

struct Args {
    const void* srcx;
    const void* srcy;
    void* y;
    int B;
    const double* post;
    int nch;
};

template <typename T, int NEM, int PPL, bool FULL>
hipError_t launch_t(const DevPlan& P, const Args& a, hipStream_t s) {
    const int ngx = (P.nx + 3) / 4, ngy = (P.ny + 3) / 4;
    const long long nw = (long long)(ngx + ngy) * a.B * a.nch;
    if (nw > (1LL << 30)) return hipErrorInvalidValue;  // grid size (one block per work item)
    const int nwork = (int)nw;
    const size_t lds = fwd_w_block_lds(line_max(P));
    auto kern = k_fwd_w<T, NEM, PPL, FULL>;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(grid_for(nwork)), dim3(NT), lds, s, P, (const T*)a.srcx, (const T*)a.srcy, (T*)a.y,
                       a.B, ngx, ngy, a.post, a.nch);
    return hipGetLastError();
}

// Every swept line has exactly n pixels (rows of the X-axis views, columns of the Y-axis views).
bool all_lines(const DevPlan& P, int n) { return (P.nx == 0 || P.W == n) && (P.ny == 0 || P.H == n); }

template <typename T, int NEM, int PPL>
hipError_t by_full(const DevPlan& P, const Args& a, hipStream_t s) {
    if (all_lines(P, 64 * PPL)) return launch_t<T, NEM, PPL, true>(P, a, s);
    return launch_t<T, NEM, PPL, false>(P, a, s);
}

// Pixels per lane: 64 * PPL >= n.
template <typename T, int NEM>
hipError_t by_ppl(const DevPlan& P, const Args& a, hipStream_t s) {
    const int n = line_max(P);
    if (n <= 256) return by_full<T, NEM, 4>(P, a, s);
    if (n <= 512) return by_full<T, NEM, 8>(P, a, s);
    if (n <= 1024) return by_full<T, NEM, 16>(P, a, s);
    return hipErrorInvalidValue;
}

// Edge slots of 64 per view: NEM >= ceil((D+1)/64).
template <typename T>
hipError_t by_nem(const DevPlan& P, const Args& a, hipStream_t s) {
    const int ne = (P.D + 1 + 63) / 64;
    if (ne <= 4) return by_ppl<T, 4>(P, a, s);
    if (ne <= 6) return by_ppl<T, 6>(P, a, s);
    if (ne <= 8) return by_ppl<T, 8>(P, a, s);
    if (ne <= 12) return by_ppl<T, 12>(P, a, s);
    if (ne <= 14) return by_ppl<T, 14>(P, a, s);
    if (ne <= 16) return by_ppl<T, 16>(P, a, s);
    if (ne <= 20) return by_ppl<T, 20>(P, a, s);
    if (ne <= 24) return by_ppl<T, 24>(P, a, s);
    if (ne <= 26) return by_ppl<T, 26>(P, a, s);  // 1536 bins: 25 slots
    if (ne <= 32) return by_ppl<T, 32>(P, a, s);
    return hipErrorInvalidValue;
}

}  // namespace

bool fwd_w_supported(const DevPlan& P) { return line_max(P) <= 1024 && P.D <= 2046; }

hipError_t launch_fwd_w(const DevPlan& P, int dtype, const void* srcx, const void* srcy, void* y, int B,
                        const double* post, int nch, hipStream_t s) {
    if (B <= 0 || P.nx + P.ny == 0) return hipSuccess;
    if (!fwd_w_supported(P)) return hipErrorInvalidValue;
    if (nch < 1) nch = 1;
    if ((P.nx && nch > P.H) || (P.ny && nch > P.W)) return hipErrorInvalidValue;
    Args a{srcx, srcy, y, B, post, nch};
    if (dtype == DT_F32) return by_nem<float>(P, a, s);
    return by_nem<double>(P, a, s);
}

}  // namespace ctlib
