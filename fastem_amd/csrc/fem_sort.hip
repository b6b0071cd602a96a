// fem_sort.hip -- the slot order of the pruned FP64 E-step: the points sorted
// by the component of their largest log-density in the last E-step, so that a
// workgroup's points share few components and the E-step's bound skips the
// others for whole waves (fem_tiles.h estep64_unit).  The sort is a stable
// radix sort of (component, slot) pairs (hipCUB, deterministic), the points'
// rows are gathered from the original X into Xs, and the permutation is kept
// so that posteriors and classes leave the context in point order.  No
// reference counterpart: the reference evaluates every component for every
// point (EM2/MatrixEM.cpp:69-87).
#include <hipcub/hipcub.hpp>

#include "fem_internal.h"

namespace fem {

__global__ void k_iota(int *__restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int)i;
}

// New slot s takes the point of old slot rel[s] (perm: old slot -> point, or
// the identity) and its row of X.  One thread per (slot, 4 features).
__global__ void k_apply_order(const int *__restrict__ rel, const int *__restrict__ perm, int *__restrict__ perm_out,
                              const float *__restrict__ X, float *__restrict__ Xs, int64_t npad, int dpad) {
    const int q = dpad / 4;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad * q) return;
    const int64_t s = i / q;
    const int f = (int)(i % q);
    const int p = perm ? perm[rel[s]] : rel[s];
    if (f == 0) perm_out[s] = p;
    reinterpret_cast<float4 *>(Xs + s * dpad)[f] = reinterpret_cast<const float4 *>(X + (int64_t)p * dpad)[f];
}

static hipError_t alloc_sort(fem_ctx *c) {
    const size_t np = (size_t)c->npad;
    hipError_t e = hipMalloc((void **)&c->Xs, np * c->dpad * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void **)&c->perm, np * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->perm2, np * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->ref2, np * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->slots, np * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void **)&c->rel, np * sizeof(int));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, c->slots, (int64_t)np);
    return hipGetLastError();
}

// Key bits of 0..K (K: padding rows, sorted last).
static int key_bits(int K) {
    int bits = 1;
    while ((1 << bits) <= K) bits++;
    return bits;
}

// Called before a pruned FP64 E-step: re-sorts the slots on the components
// the last pruned E-step named (after 1, 2, 4, ... E-steps since the
// parameters were set, then every 16), else leaves the order.
hipError_t sort_points(fem_ctx *c) {
    if (!c->refs_valid) return hipSuccess;
    if (++c->esteps_since_sort < c->sort_every) return hipSuccess;
    if (!c->Xs)
        if (hipError_t e = alloc_sort(c)) return e;
    const int bits = key_bits(c->K);
    // the temporary storage for exactly this call's arguments (the sort's
    // passes, and so its scratch, depend on the key bits and the count)
    size_t need = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, reinterpret_cast<const unsigned *>(c->ref),
                                                      reinterpret_cast<unsigned *>(c->ref2), c->slots, c->rel,
                                                      (int)c->npad, 0, bits, c->stream);
    if (e != hipSuccess) return e;
    if (need > c->sort_tmp_bytes) {
        if (c->sort_tmp) (void)hipFree(c->sort_tmp);
        c->sort_tmp = nullptr;
        c->sort_tmp_bytes = 0;
        if ((e = hipMalloc(&c->sort_tmp, need)) != hipSuccess) return e;
        c->sort_tmp_bytes = need;
    }
    e = hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, c->sort_tmp_bytes,
                                                      reinterpret_cast<const unsigned *>(c->ref),
                                                      reinterpret_cast<unsigned *>(c->ref2), c->slots, c->rel,
                                                      (int)c->npad, 0, bits, c->stream);
    if (e != hipSuccess) return e;
    const int64_t work = c->npad * (c->dpad / 4);
    hipLaunchKernelGGL(k_apply_order, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, c->stream, c->rel,
                       c->sorted ? c->perm : (const int *)nullptr, c->perm2, c->X, c->Xs, c->npad, c->dpad);
    std::swap(c->perm, c->perm2);
    std::swap(c->ref, c->ref2);
    c->sorted = true;
    c->ref_ok = false;  // the BF16x3 references are per slot
    c->esteps_since_sort = 0;
    c->sort_every = c->sort_every < 16 ? 2 * c->sort_every : 16;
    return hipGetLastError();
}

// Back to point order (new data, k-means passes that write posteriors per
// point): the slots are the points again and their components are unknown.
void reset_order(fem_ctx *c) {
    c->sorted = false;
    c->refs_valid = false;
    c->esteps_since_sort = 0;
    c->sort_every = 1;
}

// Row i of the point-ordered output of slot s: perm[s], or s before a sort.
__global__ void k_unpermute_int(const int *__restrict__ in, const int *__restrict__ perm, int64_t n,
                                int *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[perm[s]] = in[s];
}

hipError_t launch_unpermute_int(const fem_ctx *c, const int *in, int *out) {
    hipLaunchKernelGGL(k_unpermute_int, dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0, c->stream, in, c->perm,
                       c->n, out);
    return hipGetLastError();
}

}  // namespace fem
