// fem_setnet.hip -- the learned initializer's set network on the GPU: FP32
// forward, hand-written backward and Adam for training, and an inference
// forward whose products run on BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulation).
//
// Per point set b of n normalised points x_i in R^d (width W = 64, K outputs):
//   h1 = gelu(x W1 + b1), h2 = gelu(h1 W2 + b2)        per point
//   g = mean_i h2_i;  q_k = S_k + g Wg                 K seed queries
//   a_ik = softmax_i(q_k . h2_i / 8);  z_k = sum_i a_ik h2_i
//   mu_k = z_k Wm + bm;  lv_k = z_k Wv + bv            means, log-variances
//   m_i = sum_j softmax_j((h2_i Wp) . (h2_j Wq) / 8) x_j   point head
// Loss (per set, against a fitted mixture M_j, V_j, p_j):
//   sum_j p_j |mu_k(j) - M_j|^2 + (1/K) sum_k min_j |mu_k - M_j|^2
//   + lam sum_j p_j |lv_k(j) - V_j|^2 / d,  k(j) = argmin_k |mu_k - M_j|^2
//   + gam (1/n) sum_i |m_i - T_i|^2        (T_i: mean of point i's cluster).
// Training arithmetic is FP32; every reduction has a fixed order.  Parameters
// are one flat FP32 vector: W1[d][64] b1[64] W2[64][64] b2[64] S[K][64]
// Wg[64][64] Wm[64][d] bm[d] Wv[64][d] bv[d] Wp[64][64] Wq[64][64].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/fastem_abi.h"
#include "../../include/fastem_setnet.h"

namespace {

constexpr int W = 64;

struct Offsets {
    int W1, b1, W2, b2, S, Wg, Wm, bm, Wv, bv, Wp, Wq, total;
};

Offsets offsets(int d, int K) {
    Offsets o;
    int p = 0;
    o.W1 = p; p += d * W;
    o.b1 = p; p += W;
    o.W2 = p; p += W * W;
    o.b2 = p; p += W;
    o.S = p; p += K * W;
    o.Wg = p; p += W * W;
    o.Wm = p; p += W * d;
    o.bm = p; p += d;
    o.Wv = p; p += W * d;
    o.bv = p; p += d;
    o.Wp = p; p += W * W;
    o.Wq = p; p += W * W;
    o.total = p;
    return o;
}

__device__ inline float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_g(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
}

// Z = X Wt + b, H = gelu(Z); X: M x I, Wt: I x 64.  64 rows per workgroup,
// thread t computes row t/4, columns 16*(t%4) .. +15.
template <int I>
__global__ __launch_bounds__(256) void k_dense_fwd(const float *__restrict__ X, int M, const float *__restrict__ Wt,
                                                   const float *__restrict__ b, float *__restrict__ Z,
                                                   float *__restrict__ H) {
    __shared__ float xs[64][I + 1];
    __shared__ float ws[I][W];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * 64;
    for (int e = t; e < 64 * I; e += 256) {
        const int r = e / I, i = e % I;
        xs[r][i] = (r0 + r < M) ? X[(size_t)(r0 + r) * I + i] : 0.f;
    }
    for (int e = t; e < I * W; e += 256) ws[e / W][e % W] = Wt[e];
    __syncthreads();
    const int r = t >> 2, c0 = (t & 3) * 16;
    if (r0 + r >= M) return;
    for (int c = c0; c < c0 + 16; c++) {
        float s = b[c];
        for (int i = 0; i < I; i++) s += xs[r][i] * ws[i][c];
        Z[(size_t)(r0 + r) * W + c] = s;
        H[(size_t)(r0 + r) * W + c] = gelu_f(s);
    }
}

// g[b] = mean over the n points of H2 rows of set b.
__global__ void k_pool(const float *__restrict__ H2, int n, float *__restrict__ g) {
    const int b = blockIdx.x, w = threadIdx.x;
    float s = 0.f;
    for (int i = 0; i < n; i++) s += H2[((size_t)b * n + i) * W + w];
    g[b * W + w] = s / (float)n;
}

// q[b][k] = S[k] + g[b] Wg
__global__ void k_query(const float *__restrict__ g, const float *__restrict__ S, const float *__restrict__ Wg, int K,
                        float *__restrict__ q) {
    const int b = blockIdx.x;
    for (int e = threadIdx.x; e < K * W; e += blockDim.x) {
        const int k = e / W, w = e % W;
        float s = S[k * W + w];
        for (int v = 0; v < W; v++) s += g[b * W + v] * Wg[v * W + w];
        q[((size_t)b * K + k) * W + w] = s;
    }
}

// Attention of query (b,k) over the n points: alpha (stored) and z.
__global__ __launch_bounds__(256) void k_attn_fwd(const float *__restrict__ H2, const float *__restrict__ q, int n,
                                                  int K, float *__restrict__ alpha, float *__restrict__ z) {
    __shared__ float qs[W];
    __shared__ float red[256];
    __shared__ float zp[4][W];
    const int bk = blockIdx.x, b = bk / K, t = threadIdx.x;
    if (t < W) qs[t] = q[(size_t)bk * W + t];
    __syncthreads();
    float *al = alpha + (size_t)bk * n;
    float mx = -INFINITY;
    for (int i = t; i < n; i += 256) {
        const float *h = H2 + ((size_t)b * n + i) * W;
        float s = 0.f;
        for (int w = 0; w < W; w++) s += qs[w] * h[w];
        s *= 0.125f;
        al[i] = s;
        mx = fmaxf(mx, s);
    }
    red[t] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] = fmaxf(red[t], red[t + st]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.f;
    for (int i = t; i < n; i += 256) {
        const float e = expf(al[i] - mx);
        al[i] = e;
        se += e;
    }
    red[t] = se;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    const float inv = 1.0f / red[0];
    __syncthreads();
    for (int i = t; i < n; i += 256) al[i] *= inv;
    __syncthreads();
    // z = sum_i alpha_i h2_i: 4 point groups x 64 features, combined in order
    const int grp = t >> 6, w = t & 63;
    float s = 0.f;
    for (int i = grp; i < n; i += 4) s += al[i] * H2[((size_t)b * n + i) * W + w];
    zp[grp][w] = s;
    __syncthreads();
    if (t < W) z[(size_t)bk * W + t] = (zp[0][t] + zp[1][t]) + (zp[2][t] + zp[3][t]);
}

// mu = z Wm + bm, lv = z Wv + bv for the B*K rows.
__global__ void k_heads(const float *__restrict__ z, int rows, int d, const float *__restrict__ Wm,
                        const float *__restrict__ bm, const float *__restrict__ Wv, const float *__restrict__ bv,
                        float *__restrict__ mu, float *__restrict__ lv) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * d) return;
    const int r = e / d, j = e % d;
    float sm = bm[j], sv = bv[j];
    for (int w = 0; w < W; w++) {
        const float zw = z[(size_t)r * W + w];
        sm += zw * Wm[w * d + j];
        sv += zw * Wv[w * d + j];
    }
    mu[e] = sm;
    lv[e] = sv;
}

// Loss of set b and its gradient w.r.t. mu and lv (one thread per set, fixed
// order); targets M, V: J x d, weights pw: J.
__global__ void k_loss(const float *__restrict__ mu, const float *__restrict__ lv, const float *__restrict__ M,
                       const float *__restrict__ V, const float *__restrict__ pw, int B, int K, int J, int d,
                       float lam, float *__restrict__ dmu, float *__restrict__ dlv, float *__restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const float *mb = mu + (size_t)b * K * d;
    const float *vb = lv + (size_t)b * K * d;
    const float *Mb = M + (size_t)b * J * d;
    const float *Vb = V + (size_t)b * J * d;
    float *dm = dmu + (size_t)b * K * d;
    float *dv = dlv + (size_t)b * K * d;
    for (int e = 0; e < K * d; e++) {
        dm[e] = 0.f;
        dv[e] = 0.f;
    }
    auto dist = [&](int k, int j) {
        float s = 0.f;
        for (int a = 0; a < d; a++) {
            const float t = mb[k * d + a] - Mb[j * d + a];
            s += t * t;
        }
        return s;
    };
    float L = 0.f;
    for (int j = 0; j < J; j++) {
        int kb = 0;
        float best = dist(0, j);
        for (int k = 1; k < K; k++) {
            const float s = dist(k, j);
            if (s < best) {
                best = s;
                kb = k;
            }
        }
        const float p = pw[(size_t)b * J + j];
        float lvl = 0.f;
        for (int a = 0; a < d; a++) {
            const float df = mb[kb * d + a] - Mb[j * d + a];
            const float dd = vb[kb * d + a] - Vb[j * d + a];
            dm[kb * d + a] += 2.f * p * df;
            dv[kb * d + a] += 2.f * lam * p * dd / (float)d;
            lvl += dd * dd;
        }
        L += p * best + lam * p * lvl / (float)d;
    }
    for (int k = 0; k < K; k++) {
        int jb = 0;
        float best = dist(k, 0);
        for (int j = 1; j < J; j++) {
            const float s = dist(k, j);
            if (s < best) {
                best = s;
                jb = j;
            }
        }
        for (int a = 0; a < d; a++) dm[k * d + a] += 2.f * (mb[k * d + a] - Mb[jb * d + a]) / (float)K;
        L += best / (float)K;
    }
    loss[b] = L;
}

// Head backward: dz = dmu Wm^T + dlv Wv^T (rows x W).
__global__ void k_heads_bwd_data(const float *__restrict__ dmu, const float *__restrict__ dlv, int rows, int d,
                                 const float *__restrict__ Wm, const float *__restrict__ Wv, float *__restrict__ dz) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * W) return;
    const int r = e / W, w = e % W;
    float s = 0.f;
    for (int j = 0; j < d; j++) s += dmu[(size_t)r * d + j] * Wm[w * d + j] + dlv[(size_t)r * d + j] * Wv[w * d + j];
    dz[e] = s;
}

// dWm[w][j] = sum_r z[r][w] dmu[r][j], dbm[j] = sum_r dmu[r][j] (same for v).
__global__ void k_heads_bwd_weight(const float *__restrict__ z, const float *__restrict__ dmu,
                                   const float *__restrict__ dlv, int rows, int d, float *__restrict__ gWm,
                                   float *__restrict__ gbm, float *__restrict__ gWv, float *__restrict__ gbv) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < W * d) {
        const int w = e / d, j = e % d;
        float sm = 0.f, sv = 0.f;
        for (int r = 0; r < rows; r++) {
            sm += z[(size_t)r * W + w] * dmu[(size_t)r * d + j];
            sv += z[(size_t)r * W + w] * dlv[(size_t)r * d + j];
        }
        gWm[e] = sm;
        gWv[e] = sv;
    } else if (e < W * d + d) {
        const int j = e - W * d;
        float sm = 0.f, sv = 0.f;
        for (int r = 0; r < rows; r++) {
            sm += dmu[(size_t)r * d + j];
            sv += dlv[(size_t)r * d + j];
        }
        gbm[j] = sm;
        gbv[j] = sv;
    }
}

// Attention backward for query (b,k): ds_i = alpha_i (da_i - sum alpha da)/8
// with da_i = dz_k . h2_i (stored), and dq_k = sum_i ds_i h2_i.
__global__ __launch_bounds__(256) void k_attn_bwd(const float *__restrict__ H2, const float *__restrict__ alpha,
                                                  const float *__restrict__ dz, int n, int K, float *__restrict__ ds,
                                                  float *__restrict__ dq) {
    __shared__ float dzs[W];
    __shared__ float red[256];
    __shared__ float qp[4][W];
    const int bk = blockIdx.x, b = bk / K, t = threadIdx.x;
    if (t < W) dzs[t] = dz[(size_t)bk * W + t];
    __syncthreads();
    const float *al = alpha + (size_t)bk * n;
    float *dsb = ds + (size_t)bk * n;
    float acc = 0.f;
    for (int i = t; i < n; i += 256) {
        const float *h = H2 + ((size_t)b * n + i) * W;
        float s = 0.f;
        for (int w = 0; w < W; w++) s += dzs[w] * h[w];
        dsb[i] = s;
        acc += al[i] * s;
    }
    red[t] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    const float tot = red[0];
    __syncthreads();
    for (int i = t; i < n; i += 256) dsb[i] = al[i] * (dsb[i] - tot) * 0.125f;
    __syncthreads();
    const int grp = t >> 6, w = t & 63;
    float s = 0.f;
    for (int i = grp; i < n; i += 4) s += dsb[i] * H2[((size_t)b * n + i) * W + w];
    qp[grp][w] = s;
    __syncthreads();
    if (t < W) dq[(size_t)bk * W + t] = (qp[0][t] + qp[1][t]) + (qp[2][t] + qp[3][t]);
}

// Query backward: dqs[b] = sum_k dq[b][k]; dg[b] = dqs[b] Wg^T;
// dS[k] = sum_b dq[b][k]; dWg[v][w] = sum_b g[b][v] dqs[b][w].
__global__ void k_query_bwd_set(const float *__restrict__ dq, int K, const float *__restrict__ Wg,
                                float *__restrict__ dqs, float *__restrict__ dg) {
    __shared__ float s[W];
    const int b = blockIdx.x, w = threadIdx.x;
    float a = 0.f;
    for (int k = 0; k < K; k++) a += dq[((size_t)b * K + k) * W + w];
    s[w] = a;
    dqs[b * W + w] = a;
    __syncthreads();
    float t = 0.f;
    for (int u = 0; u < W; u++) t += s[u] * Wg[w * W + u];
    dg[b * W + w] = t;
}

__global__ void k_query_bwd_weight(const float *__restrict__ dq, const float *__restrict__ g,
                                   const float *__restrict__ dqs, int B, int K, float *__restrict__ gS,
                                   float *__restrict__ gWg) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < K * W) {
        const int k = e / W, w = e % W;
        float s = 0.f;
        for (int b = 0; b < B; b++) s += dq[((size_t)b * K + k) * W + w];
        gS[e] = s;
    } else if (e < K * W + W * W) {
        const int f = e - K * W, v = f / W, w = f % W;
        float s = 0.f;
        for (int b = 0; b < B; b++) s += g[b * W + v] * dqs[b * W + w];
        gWg[f] = s;
    }
}

// dZ2 row (b,i): (sum_k alpha dz_k + ds q_k + dg/n + dh2x) * gelu'(Z2), dh2x
// being the point head's contribution.
__global__ void k_dh2(const float *__restrict__ alpha, const float *__restrict__ dz, const float *__restrict__ ds,
                      const float *__restrict__ q, const float *__restrict__ dg, const float *__restrict__ dh2x,
                      const float *__restrict__ Z2, int n, int K, int rows, float *__restrict__ dZ2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * W) return;
    const int r = e / W, w = e % W, b = r / n, i = r % n;
    float s = dg[b * W + w] / (float)n + dh2x[e];
    for (int k = 0; k < K; k++) {
        const size_t bk = (size_t)b * K + k;
        s += alpha[bk * n + i] * dz[bk * W + w] + ds[bk * n + i] * q[bk * W + w];
    }
    dZ2[e] = s * gelu_g(Z2[e]);
}

// dX = dZ Wt^T (M x I) and, if Zin != null, dX *= gelu'(Zin).
template <int I>
__global__ void k_dense_bwd_data(const float *__restrict__ dZ, int M, const float *__restrict__ Wt,
                                 const float *__restrict__ Zin, float *__restrict__ dX) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= M * I) return;
    const int m = e / I, i = e % I;
    float s = 0.f;
    for (int o = 0; o < W; o++) s += dZ[(size_t)m * W + o] * Wt[i * W + o];
    dX[e] = Zin ? s * gelu_g(Zin[e]) : s;
}

// Partial weight gradients over row chunks of 256: part[c][i][o] = sum over
// the chunk of X[m][i] dZ[m][o]; part[c][I][o] = sum dZ[m][o].
template <int I>
__global__ __launch_bounds__(256) void k_dense_bwd_weight_part(const float *__restrict__ X,
                                                               const float *__restrict__ dZ, int M,
                                                               float *__restrict__ part) {
    const int c = blockIdx.x, t = threadIdx.x;
    const int m0 = c * 256, m1 = (m0 + 256 < M) ? m0 + 256 : M;
    for (int e = t; e < (I + 1) * W; e += 256) {
        const int i = e / W, o = e % W;
        float s = 0.f;
        for (int m = m0; m < m1; m++) s += ((i < I) ? X[(size_t)m * I + i] : 1.0f) * dZ[(size_t)m * W + o];
        part[(size_t)c * (I + 1) * W + e] = s;
    }
}

// Sum the partials (fixed order) into gW (I x W) and gb (W).
__global__ void k_sum_parts(const float *__restrict__ part, int chunks, int I, float *__restrict__ gW,
                            float *__restrict__ gb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (I + 1) * W) return;
    float s = 0.f;
    for (int c = 0; c < chunks; c++) s += part[(size_t)c * (I + 1) * W + e];
    if (e < I * W) gW[e] = s;
    else gb[e - I * W] = s;
}

__global__ void k_adam(float *__restrict__ th, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float gscale) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float g = gr[e] * gscale;
    const float mm = b1 * m[e] + (1.f - b1) * g;
    const float vv = b2 * v[e] + (1.f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    th[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ---------------------------------------------------------------------------
// Point head: P = H2 Wp, Q = H2 Wq, E_ij = P_i . Q_j / 8, beta = softmax_j(E),
// m_i = sum_j beta_ij x_j.

// A = X Wa, Bo = X Wb (no bias, no activation); X: M x 64.
__global__ __launch_bounds__(256) void k_linear2(const float *__restrict__ X, int M, const float *__restrict__ Wa,
                                                 const float *__restrict__ Wb, float *__restrict__ A,
                                                 float *__restrict__ Bo) {
    __shared__ float xs[64][W + 1];
    __shared__ float wa[W][W], wb[W][W];
    const int t = threadIdx.x;
    const int r0 = blockIdx.x * 64;
    for (int e = t; e < 64 * W; e += 256) {
        const int r = e / W, i = e % W;
        xs[r][i] = (r0 + r < M) ? X[(size_t)(r0 + r) * W + i] : 0.f;
        wa[e / W][e % W] = Wa[e];
        wb[e / W][e % W] = Wb[e];
    }
    __syncthreads();
    const int r = t >> 2, c0 = (t & 3) * 16;
    if (r0 + r >= M) return;
    for (int c = c0; c < c0 + 16; c++) {
        float sa = 0.f, sb = 0.f;
        for (int i = 0; i < W; i++) {
            sa += xs[r][i] * wa[i][c];
            sb += xs[r][i] * wb[i][c];
        }
        A[(size_t)(r0 + r) * W + c] = sa;
        Bo[(size_t)(r0 + r) * W + c] = sb;
    }
}

// out[b][i][j] = scale * (A_i . B_j) over I features, for i, j < n (A, B:
// B x n x I); one 64 x 64 tile per workgroup.  Scores E = P Q^T / 8 (I = 64)
// and DB = dM X^T (I = d).
template <int I>
__global__ __launch_bounds__(256) void k_gemm_nt(const float *__restrict__ A, const float *__restrict__ Bm, int n,
                                                 float scale, float *__restrict__ out) {
    __shared__ float as[64][I + 1], bs[64][I + 1];
    const int t = threadIdx.x;
    const int b = blockIdx.z, i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    for (int e = t; e < 64 * I; e += 256) {
        const int r = e / I, w = e % I;
        as[r][w] = (i0 + r < n) ? A[((size_t)b * n + i0 + r) * I + w] : 0.f;
        bs[r][w] = (j0 + r < n) ? Bm[((size_t)b * n + j0 + r) * I + w] : 0.f;
    }
    __syncthreads();
    const int ti = t >> 4, tj = t & 15;
    float acc[4][4] = {};
    for (int w = 0; w < I; w++) {
        float qv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) qv[c] = bs[tj + 16 * c][w];
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const float pv = as[ti + 16 * a][w];
#pragma unroll
            for (int c = 0; c < 4; c++) acc[a][c] += pv * qv[c];
        }
    }
#pragma unroll
    for (int a = 0; a < 4; a++) {
        const int i = i0 + ti + 16 * a;
        if (i >= n) continue;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int j = j0 + tj + 16 * c;
            if (j < n) out[((size_t)b * n + i) * n + j] = acc[a][c] * scale;
        }
    }
}

// Row (b, i): softmax over j of E in place.
__global__ __launch_bounds__(256) void k_psoftmax(float *__restrict__ E, int n) {
    __shared__ float red[256];
    const int r = blockIdx.x, t = threadIdx.x;
    float *e = E + (size_t)r * n;
    float mx = -INFINITY;
    for (int j = t; j < n; j += 256) mx = fmaxf(mx, e[j]);
    red[t] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] = fmaxf(red[t], red[t + st]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.f;
    for (int j = t; j < n; j += 256) {
        const float v = expf(e[j] - mx);
        e[j] = v;
        se += v;
    }
    red[t] = se;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    const float inv = 1.0f / red[0];
    for (int j = t; j < n; j += 256) e[j] *= inv;
}

// Point loss gam/n |m_i - T_i|^2 per row and its gradient.
__global__ void k_ploss(const float *__restrict__ m, const float *__restrict__ T, int rows, int n, int d,
                        float gam, float *__restrict__ dm, float *__restrict__ lrow) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float L = 0.f;
    for (int a = 0; a < d; a++) {
        const float diff = m[(size_t)r * d + a] - T[(size_t)r * d + a];
        dm[(size_t)r * d + a] = 2.f * gam * diff / (float)n;
        L += diff * diff;
    }
    lrow[r] = gam * L / (float)n;
}

// loss[b] += sum of the set's row losses (fixed order).
__global__ void k_ploss_sum(const float *__restrict__ lrow, int B, int n, float *__restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float s = 0.f;
    for (int i = 0; i < n; i++) s += lrow[(size_t)b * n + i];
    loss[b] += s;
}

__global__ void k_fillf(float *__restrict__ p, size_t cnt, float v) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) p[e] = v;
}

// Row (b, i): ds_ij = beta_ij (db_ij - sum_j beta_ij db_ij) / 8 written over
// beta, with db = dM X^T precomputed.
__global__ __launch_bounds__(256) void k_pattn_bwd_rows(float *__restrict__ E, const float *__restrict__ DB, int n) {
    __shared__ float red[256];
    const int r = blockIdx.x, t = threadIdx.x;
    float *e = E + (size_t)r * n;
    const float *db = DB + (size_t)r * n;
    float acc = 0.f;
    for (int j = t; j < n; j += 256) acc += e[j] * db[j];
    red[t] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (t < st) red[t] += red[t + st];
        __syncthreads();
    }
    const float tot = red[0];
    for (int j = t; j < n; j += 256) e[j] = e[j] * (db[j] - tot) * 0.125f;
}

// TRANS = 0: out[b][i][:] = sum_j S[b][i][j] Y[b][j][:]   (m = beta X, dP = DS Q)
// TRANS = 1: out[b][j][:] = sum_i S[b][i][j] Y[b][i][:]   (dQ = DS^T P)
// Y and out: B x n x C.  64 output rows per workgroup, C/4 columns per thread.
template <int TRANS, int C>
__global__ __launch_bounds__(256) void k_gemm_nn(const float *__restrict__ S, const float *__restrict__ Y, int n,
                                                 float *__restrict__ out) {
    constexpr int CPT = C / 4;
    __shared__ float sl[64][65];
    __shared__ float yl[64][C + 1];
    const int t = threadIdx.x;
    const int b = blockIdx.y, r0 = blockIdx.x * 64;
    const int orow = t >> 2, c0 = (t & 3) * CPT;
    float acc[CPT] = {};
    for (int k0 = 0; k0 < n; k0 += 64) {
        for (int e = t; e < 64 * 64; e += 256) {
            const int kk = e / 64, o = e % 64;
            float v = 0.f;
            if (r0 + o < n && k0 + kk < n)
                v = TRANS ? S[((size_t)b * n + k0 + kk) * n + r0 + o] : S[((size_t)b * n + r0 + o) * n + k0 + kk];
            sl[o][kk] = v;
        }
        for (int e = t; e < 64 * C; e += 256) {
            const int kk = e / C, o = e % C;
            yl[kk][o] = (k0 + kk < n) ? Y[((size_t)b * n + k0 + kk) * C + o] : 0.f;
        }
        __syncthreads();
        for (int kk = 0; kk < 64; kk++) {
            const float dv = sl[orow][kk];
#pragma unroll
            for (int c = 0; c < CPT; c++) acc[c] += dv * yl[kk][c0 + c];
        }
        __syncthreads();
    }
    if (r0 + orow < n)
#pragma unroll
        for (int c = 0; c < CPT; c++) out[((size_t)b * n + r0 + orow) * C + c0 + c] = acc[c];
}

// dH2x = dP Wp^T + dQ Wq^T (rows x 64).
__global__ void k_pback_h2(const float *__restrict__ dP, const float *__restrict__ dQ, int rows,
                           const float *__restrict__ Wp, const float *__restrict__ Wq, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * W) return;
    const int r = e / W, w = e % W;
    float s = 0.f;
    for (int o = 0; o < W; o++) s += dP[(size_t)r * W + o] * Wp[w * W + o] + dQ[(size_t)r * W + o] * Wq[w * W + o];
    out[e] = s;
}

}  // namespace

namespace {

typedef float v4f_s __attribute__((ext_vector_type(4)));
typedef float v8f_s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf_s __attribute__((ext_vector_type(8)));

// Eight floats p[k0 .. k0+7] of a row of `cols` (zero past the end), in two
// 16-byte loads when the row segment is whole and aligned.
__device__ inline v8f_s load8_row(const float *__restrict__ row, int k0, int cols) {
    v8f_s v;
    const float *p = row + k0;
    if (k0 + 8 <= cols && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        const float4 b = *reinterpret_cast<const float4 *>(p + 4);
        v = (v8f_s){a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = (k0 + e < cols) ? p[e] : 0.f;
    }
    return v;
}

// C = act(alpha A Bt^T + bias) per batch entry z (blockIdx.z), with A: M x Kd
// (row stride lda), Bt: N x Kd (row stride ldb), C: M x N (row stride ldc).
// Products on v_mfma_f32_16x16x32_bf16 with both operands rounded to bf16
// (round to nearest even), FP32 accumulation; ACT = 1 applies GELU.  Each wave
// computes 16 rows x 64 columns (four MFMA tiles sharing the A operand); a
// workgroup of four waves covers 64 x 64.  Lane (c, g) supplies row c of the
// A tile and row c of each Bt tile at k = 8g .. 8g+7 of every 32-wide step.
template <int ACT>
__global__ __launch_bounds__(256) void k_bgemm_nt(int M, int N, int Kd, const float *__restrict__ A, int lda,
                                                  int64_t sA, const float *__restrict__ Bt, int ldb, int64_t sB,
                                                  const float *__restrict__ bias, float alpha, float *__restrict__ C,
                                                  int ldc, int64_t sC) {
    const int z = blockIdx.z;
    A += z * sA;
    Bt += z * sB;
    C += z * sC;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = lane >> 4, c = lane & 15;
    const int i0 = blockIdx.y * 64 + wave * 16, j0 = blockIdx.x * 64;
    if (i0 >= M) return;
    v4f_s acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = (v4f_s){0.f, 0.f, 0.f, 0.f};
    const int ia = i0 + c;
    const float *arow = A + (size_t)(ia < M ? ia : M - 1) * lda;
    for (int k0 = 0; k0 < Kd; k0 += 32) {
        const int kk = k0 + 8 * g;
        v8f_s av = load8_row(arow, kk, Kd);
        if (ia >= M) av = (v8f_s){0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const v8bf_s ab = __builtin_convertvector(av, v8bf_s);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int jb = j0 + 16 * t + c;
            v8f_s bv = load8_row(Bt + (size_t)(jb < N ? jb : N - 1) * ldb, kk, Kd);
            if (jb >= N) bv = (v8f_s){0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ab, __builtin_convertvector(bv, v8bf_s), acc[t], 0,
                                                             0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int j = j0 + 16 * t + c;
        if (j >= N) continue;
        const float bj = bias ? bias[j] : 0.f;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = i0 + 4 * g + v;
            if (i >= M) continue;
            const float x = alpha * acc[t][v] + bj;
            C[(size_t)i * ldc + j] = ACT ? gelu_f(x) : x;
        }
    }
}

// Out[z][j][i] = In[z][i][j] for In: rows x cols per batch entry.
__global__ void k_transpose(const float *__restrict__ In, int rows, int cols, float *__restrict__ Out) {
    const int z = blockIdx.y;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)rows * cols) return;
    const int i = (int)(e / cols), j = (int)(e % cols);
    Out[(size_t)z * rows * cols + (size_t)j * rows + i] = In[(size_t)z * rows * cols + e];
}

}  // namespace

struct fem_setnet {
    int d = 0, K = 0, device = 0;
    Offsets off{};
    hipStream_t stream = nullptr;
    float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    int cap_rows = 0, cap_sets = 0;  // workspace capacity (B*n rows, B sets)
    float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr;
    float *g = nullptr, *q = nullptr, *alpha = nullptr, *z = nullptr, *mu = nullptr, *lv = nullptr;
    float *M = nullptr, *V = nullptr, *pw = nullptr, *dmu = nullptr, *dlv = nullptr, *loss = nullptr;
    float *dz = nullptr, *ds = nullptr, *dq = nullptr, *dqs = nullptr, *dg = nullptr, *dZ2 = nullptr, *dZ1 = nullptr;
    float *part = nullptr;
    // point head
    size_t cap_e = 0;  // capacity of E in floats (sum over sets of n^2)
    float *P = nullptr, *Q = nullptr, *E = nullptr, *pm = nullptr, *T = nullptr, *dm = nullptr, *lrow = nullptr;
    float *dP = nullptr, *dQ = nullptr, *dh2x = nullptr, *dummy = nullptr, *DB = nullptr;
    // BF16-MFMA inference: transposed weights (W1^T, W2^T, Wp^T, Wq^T) and X^T
    float *wt = nullptr, *XT = nullptr;
    int step = 0;
    std::string err;
};

namespace {

int serr(fem_setnet *s, int code, const std::string &m) {
    if (s) s->err = m;
    return code;
}

#define SH(s, call)                                                                                      \
    do {                                                                                                 \
        hipError_t _e = (call);                                                                          \
        if (_e != hipSuccess) return serr((s), FEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

void free_ws(fem_setnet *s) {
    float **ptrs[] = {&s->X,  &s->Z1,  &s->H1, &s->Z2, &s->H2,  &s->g,    &s->q,  &s->alpha, &s->z,
                      &s->mu, &s->lv,  &s->M,  &s->V,  &s->pw,  &s->dmu,  &s->dlv, &s->loss, &s->dz,
                      &s->ds, &s->dq,  &s->dqs, &s->dg, &s->dZ2, &s->dZ1, &s->part, &s->P,   &s->Q,
                      &s->E,  &s->pm,  &s->T,  &s->dm, &s->lrow, &s->dP,  &s->dQ,  &s->dh2x, &s->dummy, &s->DB,
                      &s->wt, &s->XT};
    for (float **p : ptrs) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    s->cap_rows = s->cap_sets = 0;
    s->cap_e = 0;
}

// Workspace for B sets of n points; target arrays are sized for J <= 4K.
int ensure_ws(fem_setnet *s, int B, int n) {
    const int rows = B * n;
    const size_t ecount = (size_t)rows * n;
    if (rows <= s->cap_rows && B <= s->cap_sets && ecount <= s->cap_e) return FEM_OK;
    free_ws(s);
    const int d = s->d, K = s->K;
    const int Jc = s->K * 4;  // target components capacity
    auto A = [&](float **p, size_t cnt) { return hipMalloc((void **)p, std::max<size_t>(cnt, 1) * sizeof(float)); };
    const int chunks = (rows + 255) / 256;
    if (A(&s->X, (size_t)rows * d) || A(&s->Z1, (size_t)rows * W) || A(&s->H1, (size_t)rows * W) ||
        A(&s->Z2, (size_t)rows * W) || A(&s->H2, (size_t)rows * W) || A(&s->g, (size_t)B * W) ||
        A(&s->q, (size_t)B * K * W) || A(&s->alpha, (size_t)B * K * n) || A(&s->z, (size_t)B * K * W) ||
        A(&s->mu, (size_t)B * K * d) || A(&s->lv, (size_t)B * K * d) || A(&s->M, (size_t)B * Jc * d) ||
        A(&s->V, (size_t)B * Jc * d) || A(&s->pw, (size_t)B * Jc) || A(&s->dmu, (size_t)B * K * d) ||
        A(&s->dlv, (size_t)B * K * d) || A(&s->loss, (size_t)B) || A(&s->dz, (size_t)B * K * W) ||
        A(&s->ds, (size_t)B * K * n) || A(&s->dq, (size_t)B * K * W) || A(&s->dqs, (size_t)B * W) ||
        A(&s->dg, (size_t)B * W) || A(&s->dZ2, (size_t)rows * W) || A(&s->dZ1, (size_t)rows * W) ||
        A(&s->part, (size_t)chunks * 65 * W) || A(&s->P, (size_t)rows * W) || A(&s->Q, (size_t)rows * W) ||
        A(&s->E, ecount) || A(&s->pm, (size_t)rows * d) || A(&s->T, (size_t)rows * d) ||
        A(&s->dm, (size_t)rows * d) || A(&s->lrow, (size_t)rows) || A(&s->dP, (size_t)rows * W) ||
        A(&s->dQ, (size_t)rows * W) || A(&s->dh2x, (size_t)rows * W) || A(&s->dummy, (size_t)W) ||
        A(&s->DB, ecount) || A(&s->wt, (size_t)W * d + 3 * W * W) || A(&s->XT, (size_t)rows * d))
        return serr(s, FEM_ERR_HIP, "setnet workspace allocation failed");
    s->cap_rows = rows;
    s->cap_sets = B;
    s->cap_e = ecount;
    return FEM_OK;
}

template <typename F>
void by_d(int d, F &&f) {
    if (d == 8) f(std::integral_constant<int, 8>{});
    else if (d == 16) f(std::integral_constant<int, 16>{});
    else if (d == 32) f(std::integral_constant<int, 32>{});
    else f(std::integral_constant<int, 64>{});
}

// forward pass on s->X (rows = B*n): fills Z1 H1 Z2 H2 g q alpha z mu lv
void forward(fem_setnet *s, int B, int n) {
    const int rows = B * n, d = s->d, K = s->K;
    const float *th = s->theta;
    by_d(d, [&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        hipLaunchKernelGGL(k_dense_fwd<I>, dim3((rows + 63) / 64), dim3(256), 0, s->stream, s->X, rows,
                           th + s->off.W1, th + s->off.b1, s->Z1, s->H1);
    });
    hipLaunchKernelGGL(k_dense_fwd<W>, dim3((rows + 63) / 64), dim3(256), 0, s->stream, s->H1, rows, th + s->off.W2,
                       th + s->off.b2, s->Z2, s->H2);
    hipLaunchKernelGGL(k_pool, dim3(B), dim3(W), 0, s->stream, s->H2, n, s->g);
    hipLaunchKernelGGL(k_query, dim3(B), dim3(256), 0, s->stream, s->g, th + s->off.S, th + s->off.Wg, K, s->q);
    hipLaunchKernelGGL(k_attn_fwd, dim3(B * K), dim3(256), 0, s->stream, s->H2, s->q, n, K, s->alpha, s->z);
    hipLaunchKernelGGL(k_heads, dim3((B * K * d + 255) / 256), dim3(256), 0, s->stream, s->z, B * K, d,
                       th + s->off.Wm, th + s->off.bm, th + s->off.Wv, th + s->off.bv, s->mu, s->lv);
    // point head
    hipLaunchKernelGGL(k_linear2, dim3((rows + 63) / 64), dim3(256), 0, s->stream, s->H2, rows, th + s->off.Wp,
                       th + s->off.Wq, s->P, s->Q);
    const int nt = (n + 63) / 64;
    hipLaunchKernelGGL(k_gemm_nt<W>, dim3(nt, nt, B), dim3(256), 0, s->stream, s->P, s->Q, n, 0.125f, s->E);
    hipLaunchKernelGGL(k_psoftmax, dim3(rows), dim3(256), 0, s->stream, s->E, n);
    by_d(d, [&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        hipLaunchKernelGGL((k_gemm_nn<0, I>), dim3(nt, B), dim3(256), 0, s->stream, s->E, s->X, n, s->pm);
    });
}

// Inference forward on s->X with the products on BF16 MFMA: the dense layers,
// the point head's projections, its scores and its weighted sum; pooling, the
// seed queries, their attention and the heads stay FP32 (K x n x 64 work).
// Fills H1 H2 g q alpha z mu lv P Q E pm.
void forward_bf16(fem_setnet *s, int B, int n) {
    const int rows = B * n, d = s->d, K = s->K;
    const float *th = s->theta;
    float *w1t = s->wt, *w2t = w1t + W * d, *wpt = w2t + W * W, *wqt = wpt + W * W;
    auto tr = [&](const float *in, int r, int c, float *out, int batch) {
        hipLaunchKernelGGL(k_transpose, dim3((r * c + 255) / 256, batch), dim3(256), 0, s->stream, in, r, c, out);
    };
    tr(th + s->off.W1, d, W, w1t, 1);
    tr(th + s->off.W2, W, W, w2t, 1);
    tr(th + s->off.Wp, W, W, wpt, 1);
    tr(th + s->off.Wq, W, W, wqt, 1);
    tr(s->X, n, d, s->XT, B);
    const dim3 gr(1, (rows + 63) / 64, 1);
    hipLaunchKernelGGL(k_bgemm_nt<1>, gr, dim3(256), 0, s->stream, rows, W, d, s->X, d, (int64_t)0, w1t, d,
                       (int64_t)0, th + s->off.b1, 1.0f, s->H1, W, (int64_t)0);
    hipLaunchKernelGGL(k_bgemm_nt<1>, gr, dim3(256), 0, s->stream, rows, W, W, s->H1, W, (int64_t)0, w2t, W,
                       (int64_t)0, th + s->off.b2, 1.0f, s->H2, W, (int64_t)0);
    hipLaunchKernelGGL(k_pool, dim3(B), dim3(W), 0, s->stream, s->H2, n, s->g);
    hipLaunchKernelGGL(k_query, dim3(B), dim3(256), 0, s->stream, s->g, th + s->off.S, th + s->off.Wg, K, s->q);
    hipLaunchKernelGGL(k_attn_fwd, dim3(B * K), dim3(256), 0, s->stream, s->H2, s->q, n, K, s->alpha, s->z);
    hipLaunchKernelGGL(k_heads, dim3((B * K * d + 255) / 256), dim3(256), 0, s->stream, s->z, B * K, d,
                       th + s->off.Wm, th + s->off.bm, th + s->off.Wv, th + s->off.bv, s->mu, s->lv);
    hipLaunchKernelGGL(k_bgemm_nt<0>, gr, dim3(256), 0, s->stream, rows, W, W, s->H2, W, (int64_t)0, wpt, W,
                       (int64_t)0, (const float *)nullptr, 1.0f, s->P, W, (int64_t)0);
    hipLaunchKernelGGL(k_bgemm_nt<0>, gr, dim3(256), 0, s->stream, rows, W, W, s->H2, W, (int64_t)0, wqt, W,
                       (int64_t)0, (const float *)nullptr, 1.0f, s->Q, W, (int64_t)0);
    const int nt = (n + 63) / 64;
    hipLaunchKernelGGL(k_bgemm_nt<0>, dim3(nt, nt, B), dim3(256), 0, s->stream, n, n, W, s->P, W,
                       (int64_t)n * W, s->Q, W, (int64_t)n * W, (const float *)nullptr, 0.125f, s->E, n,
                       (int64_t)n * n);
    hipLaunchKernelGGL(k_psoftmax, dim3(rows), dim3(256), 0, s->stream, s->E, n);
    hipLaunchKernelGGL(k_bgemm_nt<0>, dim3((d + 63) / 64, nt, B), dim3(256), 0, s->stream, n, d, n, s->E, n,
                       (int64_t)n * n, s->XT, n, (int64_t)d * n, (const float *)nullptr, 1.0f, s->pm, d,
                       (int64_t)n * d);
}

// backward from s->dmu, s->dlv into s->grad
void backward(fem_setnet *s, int B, int n) {
    const int rows = B * n, d = s->d, K = s->K;
    const float *th = s->theta;
    float *gr = s->grad;
    hipLaunchKernelGGL(k_heads_bwd_data, dim3((B * K * W + 255) / 256), dim3(256), 0, s->stream, s->dmu, s->dlv,
                       B * K, d, th + s->off.Wm, th + s->off.Wv, s->dz);
    hipLaunchKernelGGL(k_heads_bwd_weight, dim3((W * d + d + 255) / 256), dim3(256), 0, s->stream, s->z, s->dmu,
                       s->dlv, B * K, d, gr + s->off.Wm, gr + s->off.bm, gr + s->off.Wv, gr + s->off.bv);
    hipLaunchKernelGGL(k_attn_bwd, dim3(B * K), dim3(256), 0, s->stream, s->H2, s->alpha, s->dz, n, K, s->ds, s->dq);
    hipLaunchKernelGGL(k_query_bwd_set, dim3(B), dim3(W), 0, s->stream, s->dq, K, th + s->off.Wg, s->dqs, s->dg);
    hipLaunchKernelGGL(k_query_bwd_weight, dim3((K * W + W * W + 255) / 256), dim3(256), 0, s->stream, s->dq, s->g,
                       s->dqs, B, K, gr + s->off.S, gr + s->off.Wg);
    // point head: ds over E, dP, dQ, their weight gradients and dH2 share
    const int chunks = (rows + 255) / 256;
    const int nt = (n + 63) / 64;
    by_d(d, [&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        hipLaunchKernelGGL(k_gemm_nt<I>, dim3(nt, nt, B), dim3(256), 0, s->stream, s->dm, s->X, n, 1.0f, s->DB);
    });
    hipLaunchKernelGGL(k_pattn_bwd_rows, dim3(rows), dim3(256), 0, s->stream, s->E, s->DB, n);
    hipLaunchKernelGGL((k_gemm_nn<0, W>), dim3(nt, B), dim3(256), 0, s->stream, s->E, s->Q, n, s->dP);
    hipLaunchKernelGGL((k_gemm_nn<1, W>), dim3(nt, B), dim3(256), 0, s->stream, s->E, s->P, n, s->dQ);
    hipLaunchKernelGGL(k_dense_bwd_weight_part<W>, dim3(chunks), dim3(256), 0, s->stream, s->H2, s->dP, rows,
                       s->part);
    hipLaunchKernelGGL(k_sum_parts, dim3((65 * W + 255) / 256), dim3(256), 0, s->stream, s->part, chunks, W,
                       gr + s->off.Wp, s->dummy);
    hipLaunchKernelGGL(k_dense_bwd_weight_part<W>, dim3(chunks), dim3(256), 0, s->stream, s->H2, s->dQ, rows,
                       s->part);
    hipLaunchKernelGGL(k_sum_parts, dim3((65 * W + 255) / 256), dim3(256), 0, s->stream, s->part, chunks, W,
                       gr + s->off.Wq, s->dummy);
    hipLaunchKernelGGL(k_pback_h2, dim3((rows * W + 255) / 256), dim3(256), 0, s->stream, s->dP, s->dQ, rows,
                       th + s->off.Wp, th + s->off.Wq, s->dh2x);
    hipLaunchKernelGGL(k_dh2, dim3((rows * W + 255) / 256), dim3(256), 0, s->stream, s->alpha, s->dz, s->ds, s->q,
                       s->dg, s->dh2x, s->Z2, n, K, rows, s->dZ2);
    hipLaunchKernelGGL(k_dense_bwd_weight_part<W>, dim3(chunks), dim3(256), 0, s->stream, s->H1, s->dZ2, rows,
                       s->part);
    hipLaunchKernelGGL(k_sum_parts, dim3((65 * W + 255) / 256), dim3(256), 0, s->stream, s->part, chunks, W,
                       gr + s->off.W2, gr + s->off.b2);
    hipLaunchKernelGGL(k_dense_bwd_data<W>, dim3((rows * W + 255) / 256), dim3(256), 0, s->stream, s->dZ2, rows,
                       th + s->off.W2, s->Z1, s->dZ1);
    by_d(d, [&](auto Ic) {
        constexpr int I = decltype(Ic)::value;
        hipLaunchKernelGGL(k_dense_bwd_weight_part<I>, dim3(chunks), dim3(256), 0, s->stream, s->X, s->dZ1, rows,
                           s->part);
        hipLaunchKernelGGL(k_sum_parts, dim3(((I + 1) * W + 255) / 256), dim3(256), 0, s->stream, s->part, chunks, I,
                           gr + s->off.W1, gr + s->off.b1);
    });
}

}  // namespace

extern "C" {

int fem_setnet_create(fem_setnet **out, int d, int K, uint64_t seed, int device) {
    if (!out || !(d == 8 || d == 16 || d == 32 || d == 64) || K < 1 || K > 64) return FEM_ERR_ARG;
    *out = nullptr;
    fem_setnet *s = new fem_setnet();
    s->d = d;
    s->K = K;
    s->device = device;
    s->off = offsets(d, K);
    const int P = s->off.total;
    std::vector<float> th(P, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    auto fill = [&](int o, int rows, int cols, double scale) {
        for (int e = 0; e < rows * cols; e++) th[o + e] = (float)(nd(rng) * scale);
    };
    fill(s->off.W1, d, W, 1.0 / std::sqrt((double)d));
    fill(s->off.W2, W, W, 1.0 / std::sqrt((double)W));
    fill(s->off.S, K, W, 1.0);
    fill(s->off.Wg, W, W, 1.0 / std::sqrt((double)W));
    fill(s->off.Wm, W, d, 1.0 / std::sqrt((double)W));
    fill(s->off.Wv, W, d, 1.0 / std::sqrt((double)W));
    fill(s->off.Wp, W, W, 1.0 / std::sqrt((double)W));
    fill(s->off.Wq, W, W, 1.0 / std::sqrt((double)W));
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&s->theta, P * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&s->grad, P * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&s->m, P * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&s->v, P * sizeof(float)) != hipSuccess ||
        hipMemcpy(s->theta, th.data(), P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        // on the net's own (non-blocking) stream, finished before return
        hipMemsetAsync(s->m, 0, P * sizeof(float), s->stream) != hipSuccess ||
        hipMemsetAsync(s->v, 0, P * sizeof(float), s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
        delete s;
        return FEM_ERR_HIP;
    }
    *out = s;
    return FEM_OK;
}

int fem_setnet_destroy(fem_setnet *s) {
    if (!s) return FEM_OK;
    (void)hipSetDevice(s->device);
    (void)hipStreamSynchronize(s->stream);
    free_ws(s);
    (void)hipFree(s->theta);
    (void)hipFree(s->grad);
    (void)hipFree(s->m);
    (void)hipFree(s->v);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return FEM_OK;
}

const char *fem_setnet_last_error(const fem_setnet *s) { return s ? s->err.c_str() : ""; }

int fem_setnet_num_params(const fem_setnet *s) { return s ? s->off.total : 0; }

int fem_setnet_shape(const fem_setnet *s, int *d, int *K) {
    if (!s) return FEM_ERR_ARG;
    if (d) *d = s->d;
    if (K) *K = s->K;
    return FEM_OK;
}

// Parameter copies wait for the net's stream (non-blocking: the null-stream
// copies would not order against its kernels).
int fem_setnet_get_params(fem_setnet *s, float *theta) {
    if (!s || !theta) return FEM_ERR_ARG;
    SH(s, hipStreamSynchronize(s->stream));
    SH(s, hipMemcpy(theta, s->theta, sizeof(float) * s->off.total, hipMemcpyDeviceToHost));
    return FEM_OK;
}

int fem_setnet_set_params(fem_setnet *s, const float *theta) {
    if (!s || !theta) return FEM_ERR_ARG;
    SH(s, hipStreamSynchronize(s->stream));
    SH(s, hipMemcpy(s->theta, theta, sizeof(float) * s->off.total, hipMemcpyHostToDevice));
    s->step = 0;
    SH(s, hipMemsetAsync(s->m, 0, sizeof(float) * s->off.total, s->stream));
    SH(s, hipMemsetAsync(s->v, 0, sizeof(float) * s->off.total, s->stream));
    SH(s, hipStreamSynchronize(s->stream));
    return FEM_OK;
}

// Forward for B normalised point sets of n points (X: B*n*d, host): means and
// log-variances (B*K*d, host) and, if m != NULL, the point head (B*n*d).
int fem_setnet_forward(fem_setnet *s, int B, int n, const float *X, float *mu, float *lv, float *m) {
    if (!s || B < 1 || n < 1 || !X || !mu || !lv) return FEM_ERR_ARG;
    SH(s, hipSetDevice(s->device));
    if (int r = ensure_ws(s, B, n)) return r;
    SH(s, hipMemcpyAsync(s->X, X, sizeof(float) * (size_t)B * n * s->d, hipMemcpyHostToDevice, s->stream));
    forward(s, B, n);
    SH(s, hipGetLastError());
    SH(s, hipMemcpyAsync(mu, s->mu, sizeof(float) * B * s->K * s->d, hipMemcpyDeviceToHost, s->stream));
    SH(s, hipMemcpyAsync(lv, s->lv, sizeof(float) * B * s->K * s->d, hipMemcpyDeviceToHost, s->stream));
    if (m) SH(s, hipMemcpyAsync(m, s->pm, sizeof(float) * (size_t)B * n * s->d, hipMemcpyDeviceToHost, s->stream));
    SH(s, hipStreamSynchronize(s->stream));
    return FEM_OK;
}

// As fem_setnet_forward, with the products on BF16 MFMA (inference).
int fem_setnet_forward_bf16(fem_setnet *s, int B, int n, const float *X, float *mu, float *lv, float *m) {
    if (!s || B < 1 || n < 1 || !X || !mu || !lv) return FEM_ERR_ARG;
    SH(s, hipSetDevice(s->device));
    if (int r = ensure_ws(s, B, n)) return r;
    SH(s, hipMemcpyAsync(s->X, X, sizeof(float) * (size_t)B * n * s->d, hipMemcpyHostToDevice, s->stream));
    forward_bf16(s, B, n);
    SH(s, hipGetLastError());
    SH(s, hipMemcpyAsync(mu, s->mu, sizeof(float) * B * s->K * s->d, hipMemcpyDeviceToHost, s->stream));
    SH(s, hipMemcpyAsync(lv, s->lv, sizeof(float) * B * s->K * s->d, hipMemcpyDeviceToHost, s->stream));
    if (m) SH(s, hipMemcpyAsync(m, s->pm, sizeof(float) * (size_t)B * n * s->d, hipMemcpyDeviceToHost, s->stream));
    SH(s, hipStreamSynchronize(s->stream));
    return FEM_OK;
}

// Loss and gradient for a batch (targets M, V: B*J*d, weights pw: B*J, point
// targets T: B*n*d or NULL for no point loss).  grad (host, may be NULL)
// receives the parameter gradient of the summed loss; if lr > 0 an Adam step
// (beta 0.9/0.999, eps 1e-8) is applied with the gradient of the mean loss
// over the batch.  *loss_out = summed loss.
int fem_setnet_step(fem_setnet *s, int B, int n, const float *X, int J, const float *M, const float *V,
                    const float *pw, const float *T, float lam, float gam, float lr, float *grad, float *loss_out) {
    if (!s || B < 1 || n < 1 || J < 1 || J > 4 * s->K || !X || !M || !V || !pw) return FEM_ERR_ARG;
    SH(s, hipSetDevice(s->device));
    if (int r = ensure_ws(s, B, n)) return r;
    const int d = s->d, K = s->K;
    const int rows = B * n;
    SH(s, hipMemcpyAsync(s->X, X, sizeof(float) * (size_t)rows * d, hipMemcpyHostToDevice, s->stream));
    SH(s, hipMemcpyAsync(s->M, M, sizeof(float) * B * J * d, hipMemcpyHostToDevice, s->stream));
    SH(s, hipMemcpyAsync(s->V, V, sizeof(float) * B * J * d, hipMemcpyHostToDevice, s->stream));
    SH(s, hipMemcpyAsync(s->pw, pw, sizeof(float) * B * J, hipMemcpyHostToDevice, s->stream));
    if (T) SH(s, hipMemcpyAsync(s->T, T, sizeof(float) * (size_t)rows * d, hipMemcpyHostToDevice, s->stream));
    forward(s, B, n);
    hipLaunchKernelGGL(k_loss, dim3((B + 63) / 64), dim3(64), 0, s->stream, s->mu, s->lv, s->M, s->V, s->pw, B, K, J,
                       d, lam, s->dmu, s->dlv, s->loss);
    if (T) {
        hipLaunchKernelGGL(k_ploss, dim3((rows + 255) / 256), dim3(256), 0, s->stream, s->pm, s->T, rows, n, d, gam,
                           s->dm, s->lrow);
        hipLaunchKernelGGL(k_ploss_sum, dim3((B + 63) / 64), dim3(64), 0, s->stream, s->lrow, B, n, s->loss);
    } else {
        const size_t cnt = (size_t)rows * d;
        hipLaunchKernelGGL(k_fillf, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s->stream, s->dm, cnt, 0.0f);
    }
    backward(s, B, n);
    SH(s, hipGetLastError());
    std::vector<float> lb(B);
    SH(s, hipMemcpyAsync(lb.data(), s->loss, sizeof(float) * B, hipMemcpyDeviceToHost, s->stream));
    if (grad) SH(s, hipMemcpyAsync(grad, s->grad, sizeof(float) * s->off.total, hipMemcpyDeviceToHost, s->stream));
    if (lr > 0.f) {
        s->step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)s->step), c2 = 1.f - std::pow(b2, (float)s->step);
        hipLaunchKernelGGL(k_adam, dim3((s->off.total + 255) / 256), dim3(256), 0, s->stream, s->theta, s->grad, s->m,
                           s->v, s->off.total, lr, b1, b2, 1e-8f, c1, c2, 1.0f / (float)B);
        SH(s, hipGetLastError());
    }
    SH(s, hipStreamSynchronize(s->stream));
    double L = 0.0;
    for (float v : lb) L += v;
    if (loss_out) *loss_out = (float)L;
    return FEM_OK;
}

}  // extern "C"
