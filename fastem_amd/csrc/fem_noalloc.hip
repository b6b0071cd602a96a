// fem_noalloc.hip -- fem_estep_noalloc: the FP64 factorisation and E-step of
// fem_factor + fem_estep over caller-owned device buffers, for callers that
// cannot allocate device memory.  Every byte of working memory (factor
// products, a zero-padded tile of rows, the tile's component-major
// posteriors, log-likelihood partials) lives in one file-scope static device
// array; the call makes no hipMalloc/hipFree and no allocating library call.
//
// The kernels are the context path's own (launch_factor, launch_estep,
// launch_transpose_post), driven through a context value on the host stack
// whose buffer pointers point into the array and into the caller's buffers.
// Rows go through the E-step in tiles of B rows (B a multiple of 4096, the
// largest that fits the array next to the factor products, at most 65,536
// and at most n rounded up), so the kernels read only the padded tile, never
// past the caller's n rows.
// Per row the arithmetic is the context E-step's, so posteriors are equal to
// fem_estep's bit for bit; the log-likelihood adds the tiles' sums in tile
// order (deterministic; differs from fem_estep only in summation order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>

#include "fem_internal.h"

namespace {

// 64 MiB per GPU, held until the process ends.
constexpr size_t kScratchDoubles = (size_t)8 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
// Device address of `scratch` on each GPU (each holds its own copy), looked
// up once per device with that device current.
void *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];

hipError_t scratch_on(int dev, double **out) {
    std::call_once(g_scratch_once[dev], [dev] {
        g_scratch_err[dev] = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
    });
    *out = static_cast<double *>(g_scratch_ptr[dev]);
    if (g_scratch_err[dev] == hipSuccess && !*out) return hipErrorInvalidSymbol;
    return g_scratch_err[dev];
}

// Tile of rows [r0, r0 + nb) of X (row stride ldx) into a B x dpad FP32
// tile, zero beyond d and beyond nb: the layout the E-step kernels read.
__global__ __launch_bounds__(256) void k_pack_rows(const float *__restrict__ X, int64_t ldx, int64_t r0,
                                                   int64_t nb, int d, int dpad, int64_t B,
                                                   float *__restrict__ tile) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= B * dpad) return;
    const int64_t r = e / dpad;
    const int q = (int)(e % dpad);
    tile[e] = (r < nb && q < d) ? X[(r0 + r) * ldx + q] : 0.0f;
}

// *acc = (first ? 0 : *acc) + *val, one thread: tiles summed in tile order.
__global__ void k_ll_accum(double *__restrict__ acc, const double *__restrict__ val, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *acc = (first ? 0.0 : *acc) + *val;
}

size_t round256(size_t doubles) { return (doubles + 31) & ~(size_t)31; }

}  // namespace

extern "C" int fem_estep_noalloc(const float *X, int64_t n, int64_t ldx, int d, int K, const double *w,
                                 const double *mu, const double *cov, double *R, int64_t ldr, double *loglik,
                                 int *info, void *stream) {
    if (!X || !w || !mu || !cov || !R) return fem::fail_global(FEM_ERR_ARG, "fem_estep_noalloc: null pointer");
    if (n < 1 || d < 1 || K < 1 || ldx < d || ldr < K)
        return fem::fail_global(FEM_ERR_ARG, "fem_estep_noalloc: bad n, d, K or row stride");
    if (d > fem::kMaxDim || K > 128)
        return fem::fail_global(FEM_ERR_ARG, "fem_estep_noalloc: d and K are limited to 128");
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev))
        return fem::fail_global(FEM_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fem::fail_global(FEM_ERR_ARG, "fem_estep_noalloc: device index");
    double *base = nullptr;
    if (hipError_t e = scratch_on(dev, &base))
        return fem::fail_global(FEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));

    fem_ctx v;  // a view: every pointer below is borrowed
    v.device = dev;
    v.stream = static_cast<hipStream_t>(stream);
    v.d = d;
    v.NB = (d + 15) / 16;
    v.dpad = 16 * v.NB;
    v.K = K;
    v.KP = 16 * ((K + 15) / 16);
    v.prec = FEM_PREC_FP64;
    v.estep_groups = v.NB <= 2 ? 4 : 2;  // as fem_create picks
    if (const char *e = std::getenv("FASTEM_ESTEP64_GROUPS")) v.estep_groups = std::atoi(e) == 4 ? 4 : 2;
    if (v.NB > 4) v.estep_groups = 2;

    // layout of the array: small per-component vectors, factor products,
    // then the row tile and its posteriors (sized by B)
    size_t off = 0;
    auto take = [&](size_t doubles) {
        double *p = base + off;
        off += round256(doubles);
        return p;
    };
    v.coef = take(K);
    v.logdet = take(K);
    v.info = reinterpret_cast<int *>(take((K + 1) / 2));
    double *ll_acc = take(1);
    v.ll_dev = take(1);
    v.U = take((size_t)K * v.dpad * v.dpad);
    v.stage64 = take((size_t)K * fem::stage64_words(v.NB));
    const size_t fixed = off;
    int64_t B = std::min<int64_t>(65536, (n + fem::kRowAlign - 1) / fem::kRowAlign * fem::kRowAlign);
    auto tile_doubles = [&](int64_t b) {
        return round256((size_t)b / 64) + round256(((size_t)b * v.dpad + 1) / 2) + round256((size_t)v.KP * b);
    };
    while (B > fem::kRowAlign && fixed + tile_doubles(B) > kScratchDoubles) B /= 2;
    if (fixed + tile_doubles(B) > kScratchDoubles)
        return fem::fail_global(FEM_ERR_ARG, "fem_estep_noalloc: shape exceeds the 64 MiB working array");
    v.ll_part = take((size_t)B / 64);
    v.X = reinterpret_cast<float *>(take(((size_t)B * v.dpad + 1) / 2));
    v.Rt = take((size_t)v.KP * B);
    v.w = const_cast<double *>(w);
    v.mu = const_cast<double *>(mu);
    v.cov = const_cast<double *>(cov);
    v.stage32 = nullptr;

    auto hip_fail = [](const char *what, hipError_t e) {
        return fem::fail_global(FEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    if (hipError_t e = fem::launch_factor(&v)) return hip_fail("factor", e);
    int hinfo[128];
    if (hipError_t e = hipMemcpyAsync(hinfo, v.info, sizeof(int) * K, hipMemcpyDeviceToHost, v.stream))
        return hip_fail("info copy", e);
    if (hipError_t e = hipStreamSynchronize(v.stream)) return hip_fail("factor", e);
    if (info) std::copy(hinfo, hinfo + K, info);
    for (int k = 0; k < K; k++)
        if (hinfo[k] != 0)
            return fem::fail_global(FEM_ERR_NOT_PD, "Cholesky decomposition failed for component " +
                                                        std::to_string(k) + " (column " +
                                                        std::to_string(hinfo[k] - 1) + ")");
    v.npad = B;
    for (int64_t r0 = 0; r0 < n; r0 += B) {
        v.n = v.ntotal = std::min<int64_t>(B, n - r0);
        hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)((B * v.dpad + 255) / 256)), dim3(256), 0, v.stream, X, ldx,
                           r0, v.n, d, v.dpad, B, v.X);
        if (hipError_t e = hipGetLastError()) return hip_fail("pack", e);
        if (hipError_t e = fem::launch_estep(&v)) return hip_fail("estep", e);
        hipLaunchKernelGGL(k_ll_accum, dim3(1), dim3(64), 0, v.stream, ll_acc, v.ll_dev, r0 == 0 ? 1 : 0);
        if (hipError_t e = hipGetLastError()) return hip_fail("accumulate", e);
        if (hipError_t e = fem::launch_transpose_post(&v, R + r0 * ldr, ldr)) return hip_fail("posteriors", e);
    }
    double ll = 0.0;
    if (hipError_t e = hipMemcpyAsync(&ll, ll_acc, sizeof(double), hipMemcpyDeviceToHost, v.stream))
        return hip_fail("loglik copy", e);
    if (hipError_t e = hipStreamSynchronize(v.stream)) return hip_fail("estep", e);
    if (loglik) *loglik = ll;
    return FEM_OK;
}
