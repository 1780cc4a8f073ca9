// MLP surrogate kernels for gfx950: the fused MFMA evaluation (potential and force), the
// tiled bf16 MFMA GEMM and the elementwise kernels of the training step, training-set
// generation, and their launchers (declared in agx_surrogate.h).
//
// The training step (forward with stored activations, backward through every layer, weight
// gradients, Adam) is built from the GEMM below in three operand layouts:
//   forward       Z[b][o]  = sum_i H[b][i] W[o][i]        (A k-contiguous, B k-contiguous)
//   data grad     D[b][i]  = sum_o G[b][o] W[o][i]        (A k-contiguous, B n-contiguous)
//   weight grad   dW[o][i] = sum_b G[b][o] H[b][i]        (A m-contiguous, B n-contiguous,
//                                                          split over b with fp32 atomics)
// Kernels that depend on the step number read it from device memory, so that one training
// step can be captured in a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include "agx_device.h"
#include "agx_mlp.h"
#include "agx_surrogate.h"

using namespace agxd;

namespace {

typedef __bf16 bf16;

#ifndef AGX_MLP_EVAL_MODE
#define AGX_MLP_EVAL_MODE 1   // paired m-tiles with weight prefetch (agx_mlp.h mlpProduct)
#endif

// ---------------- fused evaluation ----------------
// One wavefront per block; 32 points per pass; both halves of the wavefront evaluate the
// FP64 baseline of their point and the half with g == 0 writes the result.
__global__ __launch_bounds__(64) void surrogateEvalKernel(const AgxMlpDev* __restrict__ Mp,
    const AgxMultipoleDev* __restrict__ Bp, long n, const double* __restrict__ pos,
    double* __restrict__ phi, double* __restrict__ grad)
{
    __shared__ agx_bf16x8 sbuf[AGX_MLP_LDS_VEC];
    const AgxMlpDev& M = *Mp;
    const AgxMultipoleDev& B = *Bp;
    const int lane = threadIdx.x & 63;
    for(long tile = blockIdx.x; tile * 32 < n; tile += gridDim.x) {
        long i = tile * 32 + (lane & 31);
        const bool valid = i < n;
        if(!valid) i = tile * 32;
        const double x = pos[i * 3], y = pos[i * 3 + 1], z = pos[i * 3 + 2];
        float f[4], out, df[4];
        mlpFeatures(M, x, y, z, f);
        mlpWave<AGX_MLP_EVAL_MODE>(M, f, (agx_lds_bf16x8*)sbuf, out, df);
        double pb, gbx, gby, gbz, ox, oy, oz;
        FetchVec2 fb{ gmem(reinterpret_cast<const vd2*>(B.nodes)) };
        multipoleCar(B, fb, x, y, z, pb, gbx, gby, gbz);
        mlpChain(M, x, y, z, df, ox, oy, oz);
        const double k = M.outScale, fac = 1 + k * (double)out, pk = pb * k;
        if(valid && lane < 32) {
            if(phi) phi[i] = pb * fac;
            if(grad) {
                grad[i * 3] = gbx * fac + pk * ox;
                grad[i * 3 + 1] = gby * fac + pk * oy;
                grad[i * 3 + 2] = gbz * fac + pk * oz;
            }
        }
    }
}

// the raw network on given features (float [n][4]): out[n], dout/dfeat [n][4]
__global__ __launch_bounds__(64) void mlpFeatureKernel(const AgxMlpDev* __restrict__ Mp, long n,
    const float* __restrict__ feat, float* __restrict__ out, float* __restrict__ dfeat)
{
    __shared__ agx_bf16x8 sbuf[AGX_MLP_LDS_VEC];
    const AgxMlpDev& M = *Mp;
    const int lane = threadIdx.x & 63;
    for(long tile = blockIdx.x; tile * 32 < n; tile += gridDim.x) {
        long i = tile * 32 + (lane & 31);
        const bool valid = i < n;
        if(!valid) i = tile * 32;
        float f[4], o, df[4];
        for(int c = 0; c < 4; c++) f[c] = feat[i * 4 + c];
        mlpWave<AGX_MLP_EVAL_MODE>(M, f, (agx_lds_bf16x8*)sbuf, o, df);
        if(valid && lane < 32) {
            if(out) out[i] = o;
            if(dfeat) for(int c = 0; c < 4; c++) dfeat[i * 4 + c] = df[c];
        }
    }
}

// ---------------- tiled bf16 GEMM ----------------
// C[M][N] (fp32, row stride ldc) = / += sum_k A(m,k) B(k,n) over this block's k range.
// A(m,k) = A[m*lda + k] if AK else A[k*lda + m];  B(k,n) = B[n*ldb + k] if BK else B[k*ldb + n].
// 64x64 tile per 256-thread block (2x2 wavefronts of 32x32), k-step 32; M, N multiples of 64,
// each block's k range a multiple of 32 (checked by the launcher).  mode: 0 store, 1 add,
// 2 atomic add (split-k).
constexpr int GT = 64, GK = 32, GPAD = 5;   // LDS rows padded to 5 x 16 bytes
template<bool AK, bool BK>
__global__ __launch_bounds__(256) void gemmKernel(int kPer, const bf16* __restrict__ A, long lda,
    const bf16* __restrict__ B, long ldb, float* __restrict__ C, long ldc, int mode)
{
    __shared__ agx_bf16x8 As[GT * GPAD], Bs[GT * GPAD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1, g = lane >> 5;
    const long m0 = (long)blockIdx.y * GT, n0 = (long)blockIdx.x * GT;
    const long kBeg = (long)blockIdx.z * kPer, kEnd = kBeg + kPer;
    agx_f32x16 acc;
    for(int j = 0; j < 16; j++) acc[j] = 0.f;
    for(long k0 = kBeg; k0 < kEnd; k0 += GK) {
        if constexpr(AK) {
            const int row = t >> 2, kg = t & 3;
            As[row * GPAD + kg] = *reinterpret_cast<const agx_bf16x8*>(A + (m0 + row) * lda + k0 + kg * 8);
        } else {
            const int kk = t >> 3, mg = t & 7;
            const agx_bf16x8 v = *reinterpret_cast<const agx_bf16x8*>(A + (k0 + kk) * lda + m0 + mg * 8);
            bf16* as = reinterpret_cast<bf16*>(As);
            for(int q = 0; q < 8; q++) as[((mg * 8 + q) * GPAD + (kk >> 3)) * 8 + (kk & 7)] = v[q];
        }
        if constexpr(BK) {
            const int row = t >> 2, kg = t & 3;
            Bs[row * GPAD + kg] = *reinterpret_cast<const agx_bf16x8*>(B + (n0 + row) * ldb + k0 + kg * 8);
        } else {
            const int kk = t >> 3, ng = t & 7;
            const agx_bf16x8 v = *reinterpret_cast<const agx_bf16x8*>(B + (k0 + kk) * ldb + n0 + ng * 8);
            bf16* bs = reinterpret_cast<bf16*>(Bs);
            for(int q = 0; q < 8; q++) bs[((ng * 8 + q) * GPAD + (kk >> 3)) * 8 + (kk & 7)] = v[q];
        }
        __syncthreads();
#pragma unroll
        for(int kb = 0; kb < 2; kb++) {
            const agx_bf16x8 a = As[(wm * 32 + (lane & 31)) * GPAD + 2 * kb + g];
            const agx_bf16x8 b = Bs[(wn * 32 + (lane & 31)) * GPAD + 2 * kb + g];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const long col = n0 + wn * 32 + (lane & 31);
#pragma unroll
    for(int j = 0; j < 16; j++) {
        const long row = m0 + wm * 32 + 8 * (j >> 2) + 4 * g + (j & 3);
        float* c = C + row * ldc + col;
        if(mode == 2) atomicAdd(c, acc[j]);
        else if(mode == 1) *c += acc[j];
        else *c = acc[j];
    }
}

// ---------------- training-step elementwise kernels ----------------
__device__ __forceinline__ float bfr(float v) { return (float)(bf16)v; }

__global__ void fillKernel(float* p, long n, float v)
{
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = v;
}

__device__ __forceinline__ long batchOffset(const long* step, long nbatches, int batch)
{
    return (*step % nbatches) * batch;
}

// layer 1 forward: H1 = bf16(silu(W1 f + b1)), S1 = bf16-rounded silu'
__global__ void layer1Kernel(const float* __restrict__ theta, const float* __restrict__ feat,
    const long* __restrict__ step, long nbatches, int batch, bf16* __restrict__ H, float* __restrict__ S)
{
    const long off = batchOffset(step, nbatches, batch);
    for(long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)batch * 256; e += (long)gridDim.x * blockDim.x) {
        const long b = e >> 8;
        const int o = (int)(e & 255);
        const float* f = feat + (off + b) * 4;
        const float* w = theta + AGX_P_W1 + o * 4;
        const float z = theta[AGX_P_B1 + o] + w[0] * f[0] + w[1] * f[1] + w[2] * f[2] + w[3] * f[3];
        float h, s;
        siluBoth(z, h, s);
        H[e] = (bf16)h;
        S[e] = bfr(s);
    }
}

// Z + bias -> H = bf16(silu), S = bf16-rounded silu'
__global__ void biasSiluKernel(const float* __restrict__ Z, const float* __restrict__ bias, int batch,
    bf16* __restrict__ H, float* __restrict__ S)
{
    for(long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)batch * 256; e += (long)gridDim.x * blockDim.x) {
        float h, s;
        siluBoth(Z[e] + bias[e & 255], h, s);
        H[e] = (bf16)h;
        S[e] = bfr(s);
    }
}

// output layer, loss and the start of the backward; one wavefront per sample.
//   out = w5 . H4 + b5;  e = out - target;  loss += e^2 / batch;  dout = 2 e / batch
//   G = dout * w5 * S4 (float), Gb = bf16(G);  db5 += dout  (dw5 by dw5Kernel from dout)
// Sums over the block's four samples are formed in LDS before one atomic per block.
__global__ __launch_bounds__(256) void outputLossKernel(const float* __restrict__ theta,
    const bf16* __restrict__ H4, const float* __restrict__ S4, const float* __restrict__ target,
    const long* __restrict__ step, long nbatches, int batch, float* __restrict__ G, bf16* __restrict__ Gb,
    float* __restrict__ dout, float* __restrict__ grad, float* __restrict__ lossHist)
{
    __shared__ float red[2][4];
    const long off = batchOffset(step, nbatches, batch);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long b = (long)blockIdx.x * 4 + w;
    const bool valid = b < batch;
    const long bb = valid ? b : 0;
    const float* w5 = theta + AGX_P_W5;
    float part = 0.f;
    for(int q = 0; q < 4; q++) part += w5[lane * 4 + q] * (float)H4[bb * 256 + lane * 4 + q];
    for(int sh = 32; sh >= 1; sh >>= 1) part += __shfl_xor(part, sh);
    const float out = part + theta[AGX_P_B5];
    const float e = out - target[off + bb];
    const float d = valid ? 2.f * e / (float)batch : 0.f;
    if(lane == 0) {
        red[0][w] = valid ? e * e / (float)batch : 0.f;
        red[1][w] = d;
        if(valid) dout[b] = d;
    }
    if(valid)
        for(int q = 0; q < 4; q++) {
            const long idx = b * 256 + lane * 4 + q;
            const float gv = d * w5[lane * 4 + q] * S4[idx];
            G[idx] = gv;
            Gb[idx] = (bf16)gv;
        }
    __syncthreads();
    if(threadIdx.x == 0) {
        atomicAdd(lossHist + *step, red[0][0] + red[0][1] + red[0][2] + red[0][3]);
        atomicAdd(grad + AGX_P_B5, red[1][0] + red[1][1] + red[1][2] + red[1][3]);
    }
}

// output-layer weight gradient dw5[o] += sum_b dout[b] H4[b][o] (atomic partial sums over
// row chunks, one thread per neuron)
__global__ __launch_bounds__(256) void dw5Kernel(const bf16* __restrict__ H4, const float* __restrict__ dout,
    int batch, int rowsPer, float* __restrict__ grad)
{
    const int o = threadIdx.x;
    const long r0 = (long)blockIdx.x * rowsPer;
    float s = 0.f;
    for(long r = r0; r < r0 + rowsPer && r < batch; r++) s += dout[r] * (float)H4[r * 256 + o];
    atomicAdd(grad + AGX_P_W5 + o, s);
}

// bias gradient: column sums of G [batch][256] (atomic partial sums over row chunks)
__global__ __launch_bounds__(256) void colSumKernel(const float* __restrict__ G, int batch, int rowsPer,
    float* __restrict__ db)
{
    const int o = threadIdx.x;
    const long r0 = (long)blockIdx.x * rowsPer;
    float s = 0.f;
    for(long r = r0; r < r0 + rowsPer && r < batch; r++) s += G[r * 256 + o];
    atomicAdd(db + o, s);
}

// G = D * S (float) and Gb = bf16(G)
__global__ void mulSKernel(const float* __restrict__ D, const float* __restrict__ S, int batch,
    float* __restrict__ G, bf16* __restrict__ Gb)
{
    for(long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)batch * 256; e += (long)gridDim.x * blockDim.x) {
        const float g = D[e] * S[e];
        G[e] = g;
        Gb[e] = (bf16)g;
    }
}

// layer-1 gradients: dW1[o][c] += sum_b G1[b][o] f[b][c], db1[o] += sum_b G1[b][o]
__global__ __launch_bounds__(256) void layer1GradKernel(const float* __restrict__ G, const float* __restrict__ feat,
    const long* __restrict__ step, long nbatches, int batch, int rowsPer, float* __restrict__ grad)
{
    const long off = batchOffset(step, nbatches, batch);
    const int o = threadIdx.x;
    const long r0 = (long)blockIdx.x * rowsPer;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, sb = 0.f;
    for(long r = r0; r < r0 + rowsPer && r < batch; r++) {
        const float gv = G[r * 256 + o];
        const float* f = feat + (off + r) * 4;
        s0 += gv * f[0]; s1 += gv * f[1]; s2 += gv * f[2]; s3 += gv * f[3]; sb += gv;
    }
    atomicAdd(grad + AGX_P_W1 + o * 4, s0);
    atomicAdd(grad + AGX_P_W1 + o * 4 + 1, s1);
    atomicAdd(grad + AGX_P_W1 + o * 4 + 2, s2);
    atomicAdd(grad + AGX_P_W1 + o * 4 + 3, s3);
    atomicAdd(grad + AGX_P_B1 + o, sb);
}

// Adam with linear warm-up and cosine decay of the learning rate (to 1% of its peak)
__global__ void adamKernel(float* __restrict__ theta, const float* __restrict__ grad, float* __restrict__ m1,
    float* __restrict__ m2, const long* __restrict__ step, AgxTrainHyper hp)
{
    const long t = *step;
    const float prog = fminf(1.f, (float)t / (float)hp.totalSteps);
    const float warm = fminf(1.f, (float)(t + 1) / (float)hp.warmup);
    const float lr = hp.lr * warm * (0.01f + 0.99f * 0.5f * (1.f + cosf(3.14159265f * prog)));
    const float bc1 = 1.f - powf(hp.beta1, (float)(t + 1)), bc2 = 1.f - powf(hp.beta2, (float)(t + 1));
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < AGX_MLP_NPARAM; i += (long)gridDim.x * blockDim.x) {
        const float gv = grad[i];
        const float a = hp.beta1 * m1[i] + (1.f - hp.beta1) * gv;
        const float v = hp.beta2 * m2[i] + (1.f - hp.beta2) * gv * gv;
        m1[i] = a;
        m2[i] = v;
        theta[i] -= lr * (a / bc1) / (sqrtf(v / bc2) + hp.eps);
    }
}

// bf16 row-major copies of W2..W4 for the next step's GEMMs
__global__ void packTrainKernel(const float* __restrict__ theta, bf16* __restrict__ Wb)
{
    for(long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < 3L * 65536; e += (long)gridDim.x * blockDim.x) {
        const int l = (int)(e >> 16);
        const int off = l == 0 ? AGX_P_W2 : l == 1 ? AGX_P_W3 : AGX_P_W4;
        Wb[e] = (bf16)theta[off + (e & 65535)];
    }
}

__global__ void stepIncKernel(long* step) { if(threadIdx.x == 0 && blockIdx.x == 0) *step += 1; }

// ---------------- training set ----------------
__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(unsigned long long seed, long i, int c)
{
    const unsigned long long h = mix64(mix64(seed * 0x100000001B3ull + (unsigned long long)i) + (unsigned long long)c);
    return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// points with ln r uniform in [ln rmin, ln rmax] and isotropic directions
__global__ void samplePointsKernel(long n, unsigned long long seed, double lnrmin, double lnrmax, double* pos)
{
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double r = exp(lnrmin + (lnrmax - lnrmin) * u01(seed, i, 0));
        const double cz = 2 * u01(seed, i, 1) - 1, ph = 2 * M_PI * u01(seed, i, 2);
        const double sz = sqrt(fmax(0., 1 - cz * cz));
        pos[i * 3] = r * sz * cos(ph);
        pos[i * 3 + 1] = r * sz * sin(ph);
        pos[i * 3 + 2] = r * cz;
    }
}

// relative residual q = Phi / Phi_base - 1 (stored in target), its sum of squares, features
__global__ void residualKernel(long n, const double* __restrict__ pos, const double* __restrict__ phi,
    const double* __restrict__ phib, float lnScale, float* __restrict__ feat, double* __restrict__ q,
    double* __restrict__ sumsq)
{
    double acc = 0;
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double x = pos[i * 3], y = pos[i * 3 + 1], z = pos[i * 3 + 2];
        const double r = sqrt(x * x + y * y + z * z);
        feat[i * 4] = (float)(fabs(x) / r);
        feat[i * 4 + 1] = (float)(fabs(y) / r);
        feat[i * 4 + 2] = (float)(fabs(z) / r);
        feat[i * 4 + 3] = (float)(log(r) / (double)lnScale);
        const double v = phi[i] / phib[i] - 1;
        q[i] = v;
        acc += v * v;
    }
    for(int sh = 32; sh >= 1; sh >>= 1) acc += __shfl_xor(acc, sh);
    if((threadIdx.x & 63) == 0) atomicAdd(sumsq, acc);
}

__global__ void targetKernel(long n, const double* __restrict__ q, double inv, float* __restrict__ target)
{
    for(long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        target[i] = (float)(q[i] * inv);
}

int gridFor(long n, int block, int cap = 65536)
{
    long g = (n + block - 1) / block;
    return (int)(g < 1 ? 1 : g > cap ? cap : g);
}

}  // namespace

// ---------------- launchers ----------------
hipError_t agxLaunchSurrogateEval(const AgxMlpDev* dM, const AgxMultipoleDev* dBase, long n,
    const double* pos, double* phi, double* grad, hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    const long tiles = (n + 31) / 32;
    hipLaunchKernelGGL(surrogateEvalKernel, dim3((unsigned)(tiles < 65536 ? tiles : 65536)), dim3(64), 0, stream,
        dM, dBase, n, pos, phi, grad);
    return hipGetLastError();
}

hipError_t agxLaunchMlpFeatures(const AgxMlpDev* dM, long n, const float* feat, float* out, float* dfeat,
    hipStream_t stream)
{
    if(n <= 0) return hipSuccess;
    const long tiles = (n + 31) / 32;
    hipLaunchKernelGGL(mlpFeatureKernel, dim3((unsigned)(tiles < 65536 ? tiles : 65536)), dim3(64), 0, stream,
        dM, n, feat, out, dfeat);
    return hipGetLastError();
}

hipError_t agxLaunchGemm(int M, int N, int K, int splits, bool aK, bool bK, const void* A, long lda,
    const void* B, long ldb, float* C, long ldc, int mode, hipStream_t stream)
{
    if(M % GT || N % GT || splits < 1 || K % (GK * splits)) return hipErrorInvalidValue;
    const int kPer = K / splits;
    const dim3 grid(N / GT, M / GT, splits), block(256);
    const bf16* a = static_cast<const bf16*>(A);
    const bf16* b = static_cast<const bf16*>(B);
    if(aK && bK) hipLaunchKernelGGL((gemmKernel<true, true>), grid, block, 0, stream, kPer, a, lda, b, ldb, C, ldc, mode);
    else if(aK) hipLaunchKernelGGL((gemmKernel<true, false>), grid, block, 0, stream, kPer, a, lda, b, ldb, C, ldc, mode);
    else if(bK) hipLaunchKernelGGL((gemmKernel<false, true>), grid, block, 0, stream, kPer, a, lda, b, ldb, C, ldc, mode);
    else hipLaunchKernelGGL((gemmKernel<false, false>), grid, block, 0, stream, kPer, a, lda, b, ldb, C, ldc, mode);
    return hipGetLastError();
}

hipError_t agxEnqueueTrainStep(const AgxTrainBuffers& T, const AgxTrainHyper& hp, hipStream_t s)
{
    const int B = T.batch;
    const long E = (long)B * 256;
    const int eg = gridFor(E, 256, 8192);
    const int rowsPer = 64, cg = (B + rowsPer - 1) / rowsPer;
    const int splits = B / 512 > 0 ? B / 512 : 1;
    hipError_t e;
#define AGX_TRY(x) do { x; if((e = hipGetLastError()) != hipSuccess) return e; } while(0)
    AGX_TRY(hipLaunchKernelGGL(fillKernel, dim3(gridFor(AGX_MLP_NPARAM, 256, 1024)), dim3(256), 0, s,
        T.grad, (long)AGX_MLP_NPARAM, 0.f));
    // forward
    AGX_TRY(hipLaunchKernelGGL(layer1Kernel, dim3(eg), dim3(256), 0, s, T.theta, T.feat, T.step, T.nbatches, B,
        (bf16*)T.H[0], T.S[0]));
    const int boff[3] = { AGX_P_B2, AGX_P_B3, AGX_P_B4 };
    for(int l = 0; l < 3; l++) {
        if((e = agxLaunchGemm(B, 256, 256, 1, true, true, T.H[l], 256, (const bf16*)T.Wb + l * 65536L, 256,
            T.Z, 256, 0, s)) != hipSuccess) return e;
        AGX_TRY(hipLaunchKernelGGL(biasSiluKernel, dim3(eg), dim3(256), 0, s, T.Z, T.theta + boff[l], B,
            (bf16*)T.H[l + 1], T.S[l + 1]));
    }
    AGX_TRY(hipLaunchKernelGGL(outputLossKernel, dim3((B + 3) / 4), dim3(256), 0, s, T.theta, (const bf16*)T.H[3],
        T.S[3], T.target, T.step, T.nbatches, B, T.G, (bf16*)T.Gb, T.dout, T.grad, T.lossHist));
    AGX_TRY(hipLaunchKernelGGL(dw5Kernel, dim3(cg), dim3(256), 0, s, (const bf16*)T.H[3], T.dout, B, rowsPer, T.grad));
    // backward: layers 4, 3, 2
    const int woff[3] = { AGX_P_W2, AGX_P_W3, AGX_P_W4 };
    for(int l = 2; l >= 0; l--) {
        AGX_TRY(hipLaunchKernelGGL(colSumKernel, dim3(cg), dim3(256), 0, s, T.G, B, rowsPer, T.grad + boff[l]));
        if((e = agxLaunchGemm(256, 256, B, splits, false, false, T.Gb, 256, T.H[l], 256,
            T.grad + woff[l], 256, 2, s)) != hipSuccess) return e;
        if((e = agxLaunchGemm(B, 256, 256, 1, true, false, T.Gb, 256, (const bf16*)T.Wb + l * 65536L, 256,
            T.Z, 256, 0, s)) != hipSuccess) return e;
        AGX_TRY(hipLaunchKernelGGL(mulSKernel, dim3(eg), dim3(256), 0, s, T.Z, T.S[l], B, T.G, (bf16*)T.Gb));
    }
    AGX_TRY(hipLaunchKernelGGL(layer1GradKernel, dim3(cg), dim3(256), 0, s, T.G, T.feat, T.step, T.nbatches, B,
        rowsPer, T.grad));
    if(hp.lr > 0) {
        AGX_TRY(hipLaunchKernelGGL(adamKernel, dim3(gridFor(AGX_MLP_NPARAM, 256, 1024)), dim3(256), 0, s,
            T.theta, T.grad, T.m1, T.m2, T.step, hp));
        AGX_TRY(hipLaunchKernelGGL(packTrainKernel, dim3(768), dim3(256), 0, s, T.theta, (bf16*)T.Wb));
        AGX_TRY(hipLaunchKernelGGL(stepIncKernel, dim3(1), dim3(64), 0, s, T.step));
    }
#undef AGX_TRY
    return hipSuccess;
}

hipError_t agxLaunchPackTrain(const float* theta, void* Wb, hipStream_t s)
{
    hipLaunchKernelGGL(packTrainKernel, dim3(768), dim3(256), 0, s, theta, (bf16*)Wb);
    return hipGetLastError();
}

hipError_t agxLaunchSamplePoints(long n, unsigned long long seed, double rmin, double rmax, double* pos,
    hipStream_t s)
{
    hipLaunchKernelGGL(samplePointsKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, n, seed, log(rmin), log(rmax), pos);
    return hipGetLastError();
}

hipError_t agxLaunchResidual(long n, const double* pos, const double* phi, const double* phib, float lnScale,
    float* feat, double* q, double* sumsq, hipStream_t s)
{
    hipLaunchKernelGGL(residualKernel, dim3(gridFor(n, 256, 4096)), dim3(256), 0, s, n, pos, phi, phib, lnScale,
        feat, q, sumsq);
    return hipGetLastError();
}

hipError_t agxLaunchTarget(long n, const double* q, double inv, float* target, hipStream_t s)
{
    hipLaunchKernelGGL(targetKernel, dim3(gridFor(n, 256)), dim3(256), 0, s, n, q, inv, target);
    return hipGetLastError();
}
