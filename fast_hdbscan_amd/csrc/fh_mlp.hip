// Learned search radii (see fh_mlp.h).  The network is tiny (F -> 64 -> 64
// -> 1), so training runs as three plain kernels per Adam step: a forward +
// backward pass with one thread per sample, one thread per parameter summing
// its gradient over the batch in a fixed order (deterministic), and the
// Adam update.  Inference keeps the weights in LDS, one thread per point.
#include "fh_mlp.h"

#include <math.h>

namespace fh {

namespace {

constexpr int H = kMlpHidden;

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// parameter offsets
struct Off {
    int w1, b1, w2, b2, w3, b3, P;
    __host__ __device__ explicit Off(int F) {
        w1 = 0;
        b1 = H * F;
        w2 = b1 + H;
        b2 = w2 + H * H;
        w3 = b2 + H;
        b3 = w3 + H;
        P = b3 + 1;
    }
};

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ void mlp_features_kernel(int64_t n, int d, int levels, const float *__restrict__ xp,
                                    const int2 *__restrict__ range, const float *__restrict__ box,
                                    float *__restrict__ feat) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int node = 0;
    for (int depth = 0; depth < levels - 1; ++depth) {
        const int l = 2 * node + 1;
        node = p < range[l].y ? l : l + 1;
    }
    const int2 r = range[node];
    const float *lo = box + (size_t)node * 2 * d;
    const float *hi = lo + d;
    const int F = 2 + 2 * d;
    float *f = feat + p * F;
    const float cnt = (float)(r.y - r.x);
    const float lc = logf(cnt > 1.0f ? cnt : 1.0f);
    float lv = 0.0f;
    for (int k = 0; k < d; ++k) {
        const float e = hi[k] - lo[k];
        float le = logf(e + 1e-20f);
        le = le < -40.0f ? -40.0f : (le > 40.0f ? 40.0f : le);
        lv += le;
        f[2 + k] = le;
        const float x = xp[p * d + k];
        f[2 + d + k] = e > 0.0f ? (x - lo[k]) / e - 0.5f : 0.0f;
    }
    f[0] = lc;
    f[1] = (lv - lc) / (float)d;
}

__global__ void mlp_standardise_kernel(int count, int F, const float *__restrict__ feat, const int *__restrict__ rows,
                                       float *mu, float *sd) {
    __shared__ double s1[256], s2[256];
    const int j = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int64_t r = rows ? rows[i] : i;
        const double v = feat[r * F + j];
        a += v;
        b += v * v;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = count > 0 ? s1[0] / count : 0.0;
        const double var = count > 0 ? s2[0] / count - m * m : 0.0;
        mu[j] = (float)m;
        sd[j] = (float)sqrt(var > 0.0 ? var : 0.0) + 1e-6f;
    }
}

__global__ void mlp_init_kernel(MlpParams p, uint32_t seed) {
    const Off o(p.F);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= o.P) return;
    float fan = 0.0f;
    if (q < o.b1) fan = (float)p.F;
    else if (q >= o.w2 && q < o.b2) fan = (float)H;
    else if (q >= o.w3 && q < o.b3) fan = (float)H;
    float w = 0.0f;
    if (fan > 0.0f) {
        const float u = (mix32(seed * 0x9E3779B9U + (uint32_t)q) >> 8) * (1.0f / 16777216.0f);
        const float a = sqrtf(6.0f / fan);  // He-uniform
        w = (2.0f * u - 1.0f) * a;
    }
    p.w[q] = w;
    p.m[q] = 0.0f;
    p.v[q] = 0.0f;
    p.g[q] = 0.0f;
}

// Scratch layout for a batch of B samples (field-major):
//   X [F][B], H1 [H][B], H2 [H][B], D1 [H][B], D2 [H][B], D3 [B], Y [B]
struct Scratch {
    float *X, *H1, *H2, *D1, *D2, *D3, *Y;
    __device__ Scratch(float *s, int F, int B) {
        X = s;
        H1 = X + (size_t)F * B;
        H2 = H1 + (size_t)H * B;
        D1 = H2 + (size_t)H * B;
        D2 = D1 + (size_t)H * B;
        D3 = D2 + (size_t)H * B;
        Y = D3 + B;
    }
};

// forward + backward of one sample per thread; loss = mean (y - t)^2
__global__ void mlp_fwd_bwd_kernel(MlpParams p, int ntrain, int B, int off, const float *__restrict__ feat,
                                   const int *__restrict__ rows, const float *__restrict__ tgt, float *scratch,
                                   int standardise) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const int F = p.F;
    const Off o(F);
    Scratch s(scratch, F, B);
    const int si = (off + i) % ntrain;
    const int64_t r = rows ? rows[si] : si;
    const float *w = p.w;
    for (int j = 0; j < F; ++j) {
        const float v = feat[r * F + j];
        s.X[(size_t)j * B + i] = standardise ? (v - p.mu[j]) / p.sd[j] : v;
    }
    for (int u = 0; u < H; ++u) {
        float a = w[o.b1 + u];
        for (int j = 0; j < F; ++j) a = __builtin_fmaf(w[o.w1 + u * F + j], s.X[(size_t)j * B + i], a);
        s.H1[(size_t)u * B + i] = a > 0.0f ? a : 0.0f;
    }
    float y = w[o.b3];
    for (int u = 0; u < H; ++u) {
        float a = w[o.b2 + u];
        for (int j = 0; j < H; ++j) a = __builtin_fmaf(w[o.w2 + u * H + j], s.H1[(size_t)j * B + i], a);
        const float h = a > 0.0f ? a : 0.0f;
        s.H2[(size_t)u * B + i] = h;
        y = __builtin_fmaf(w[o.w3 + u], h, y);
    }
    s.Y[i] = y;
    const float d3 = 2.0f * (y - tgt[si]) / (float)B;
    s.D3[i] = d3;
    for (int u = 0; u < H; ++u) s.D2[(size_t)u * B + i] = s.H2[(size_t)u * B + i] > 0.0f ? w[o.w3 + u] * d3 : 0.0f;
    for (int j = 0; j < H; ++j) {
        float a = 0.0f;
        for (int u = 0; u < H; ++u) a = __builtin_fmaf(w[o.w2 + u * H + j], s.D2[(size_t)u * B + i], a);
        s.D1[(size_t)j * B + i] = s.H1[(size_t)j * B + i] > 0.0f ? a : 0.0f;
    }
}

// one thread per parameter: its gradient summed over the batch in order
__global__ void mlp_grad_kernel(MlpParams p, int B, const float *scratch) {
    const int F = p.F;
    const Off o(F);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= o.P) return;
    Scratch s(const_cast<float *>(scratch), F, B);
    const float *a = nullptr, *b = nullptr;  // gradient = sum_i a[i] * (b ? b[i] : 1)
    if (q < o.b1) {
        a = s.D1 + (size_t)(q / F) * B;
        b = s.X + (size_t)(q % F) * B;
    } else if (q < o.w2) {
        a = s.D1 + (size_t)(q - o.b1) * B;
    } else if (q < o.b2) {
        const int t = q - o.w2;
        a = s.D2 + (size_t)(t / H) * B;
        b = s.H1 + (size_t)(t % H) * B;
    } else if (q < o.w3) {
        a = s.D2 + (size_t)(q - o.b2) * B;
    } else if (q < o.b3) {
        a = s.D3;
        b = s.H2 + (size_t)(q - o.w3) * B;
    } else {
        a = s.D3;
    }
    float g = 0.0f;
    if (b) {
        for (int i = 0; i < B; ++i) g = __builtin_fmaf(a[i], b[i], g);
    } else {
        for (int i = 0; i < B; ++i) g += a[i];
    }
    p.g[q] = g;
}

__global__ void mlp_adam_kernel(MlpParams p, float lr, int step) {
    const Off o(p.F);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= o.P) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float g = p.g[q];
    const float m = b1 * p.m[q] + (1.0f - b1) * g;
    const float v = b2 * p.v[q] + (1.0f - b2) * g * g;
    p.m[q] = m;
    p.v[q] = v;
    const float mh = m / (1.0f - powf(b1, (float)step));
    const float vh = v / (1.0f - powf(b2, (float)step));
    p.w[q] -= lr * mh / (sqrtf(vh) + eps);
}

// inference: weights in LDS, one thread per row
__global__ __launch_bounds__(256) void mlp_predict_kernel(MlpParams p, int64_t count, const float *__restrict__ feat,
                                                          const int *__restrict__ rows, float *__restrict__ out) {
    __shared__ float sw[H * kMlpMaxF + H + H * H + H + H + 1];
    __shared__ float smu[kMlpMaxF], ssd[kMlpMaxF];
    const int F = p.F;
    const Off o(F);
    for (int q = threadIdx.x; q < o.P; q += blockDim.x) sw[q] = p.w[q];
    if (threadIdx.x < F) {
        smu[threadIdx.x] = p.mu[threadIdx.x];
        ssd[threadIdx.x] = p.sd[threadIdx.x];
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int64_t r = rows ? rows[i] : i;
    float x[kMlpMaxF];
#pragma unroll
    for (int j = 0; j < kMlpMaxF; ++j) x[j] = j < F ? (feat[r * F + j] - smu[j]) / ssd[j] : 0.0f;
    float h1[H];
#pragma unroll
    for (int u = 0; u < H; ++u) {
        float a = sw[o.b1 + u];
#pragma unroll
        for (int j = 0; j < kMlpMaxF; ++j)
            if (j < F) a = __builtin_fmaf(sw[o.w1 + u * F + j], x[j], a);
        h1[u] = a > 0.0f ? a : 0.0f;
    }
    float y = sw[o.b3];
    for (int u = 0; u < H; ++u) {
        float a = sw[o.b2 + u];
#pragma unroll
        for (int j = 0; j < H; ++j) a = __builtin_fmaf(sw[o.w2 + u * H + j], h1[j], a);
        y = __builtin_fmaf(sw[o.w3 + u], a > 0.0f ? a : 0.0f, y);
    }
    out[i] = y;
}

__global__ void mlp_residual_kernel(int ntrain, const float *pred, const float *tgt, float *stats) {
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < ntrain; i += blockDim.x) {
        const double r = (double)tgt[i] - (double)pred[i];
        a += r;
        b += r * r;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            s1[threadIdx.x] += s1[threadIdx.x + o];
            s2[threadIdx.x] += s2[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = ntrain > 0 ? s1[0] / ntrain : 0.0;
        const double var = ntrain > 0 ? s2[0] / ntrain - m * m : 0.0;
        stats[0] = (float)m;
        stats[1] = (float)sqrt(var > 0.0 ? var : 0.0);
    }
}

__global__ void mlp_log_targets_kernel(int ntrain, const int *rows, const float *core_p, float *tgt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntrain) return;
    const float c = core_p[rows[i]];
    tgt[i] = 0.5f * logf(c > 1e-30f ? c : 1e-30f);
}

__global__ void mlp_bounds_kernel(int64_t n, const float *pred, const float *stats, float z, float *bound_sq) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float lr = pred[i] + stats[0] + z * stats[1];
    const float b = expf(2.0f * lr);
    bound_sq[i] = isfinite(b) ? b : __builtin_huge_valf();
}

}  // namespace

hipError_t launch_mlp_features(int64_t n, int d, int levels, const float *xp, const int2 *range, const float *box,
                               float *feat, hipStream_t st) {
    if (2 + 2 * d > kMlpMaxF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_features_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, d, levels, xp, range, box, feat);
    return hipGetLastError();
}

hipError_t launch_mlp_standardise(int count, int F, const float *feat, const int *rows, float *mu, float *sd,
                                  hipStream_t st) {
    hipLaunchKernelGGL(mlp_standardise_kernel, dim3(F), dim3(256), 0, st, count, F, feat, rows, mu, sd);
    return hipGetLastError();
}

hipError_t launch_mlp_init(const MlpParams &p, uint32_t seed, hipStream_t st) {
    const int P = mlp_param_count(p.F);
    hipLaunchKernelGGL(mlp_init_kernel, dim3(nblk(P, 256)), dim3(256), 0, st, p, seed);
    return hipGetLastError();
}

hipError_t launch_mlp_train_step(const MlpParams &p, int ntrain, int B, int off, const float *feat, const int *rows,
                                 const float *tgt, float *scratch, float lr, int step, hipStream_t st) {
    const int P = mlp_param_count(p.F);
    hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3(nblk(B, 64)), dim3(64), 0, st, p, ntrain, B, off, feat, rows, tgt,
                       scratch, 1);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk(P, 256)), dim3(256), 0, st, p, B, scratch);
    if (lr > 0.0f) hipLaunchKernelGGL(mlp_adam_kernel, dim3(nblk(P, 256)), dim3(256), 0, st, p, lr, step);
    return hipGetLastError();
}

hipError_t launch_mlp_predict(const MlpParams &p, int64_t n, const float *feat, const int *rows, float *out,
                              hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_predict_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, feat, rows, out);
    return hipGetLastError();
}

hipError_t launch_mlp_residual(int ntrain, const float *pred, const float *tgt, float *stats, hipStream_t st) {
    hipLaunchKernelGGL(mlp_residual_kernel, dim3(1), dim3(256), 0, st, ntrain, pred, tgt, stats);
    return hipGetLastError();
}

hipError_t launch_mlp_log_targets(int ntrain, const int *rows, const float *core_p, float *tgt, hipStream_t st) {
    hipLaunchKernelGGL(mlp_log_targets_kernel, dim3(nblk(ntrain, 256)), dim3(256), 0, st, ntrain, rows, core_p, tgt);
    return hipGetLastError();
}

hipError_t launch_mlp_bounds(int64_t n, const float *pred, const float *stats, float z, float *bound_sq,
                             hipStream_t st) {
    hipLaunchKernelGGL(mlp_bounds_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, n, pred, stats, z, bound_sq);
    return hipGetLastError();
}

}  // namespace fh
