// Host-side declarations of the MLP surrogate launchers (agx_surrogate.hip) and the
// surrogate orbit kernel launcher (agx_orbit_surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "agx_kernels.h"
#include "agx_mlp.h"

struct AgxTrainHyper {
    float lr, beta1, beta2, eps;
    long totalSteps, warmup;
};

// device buffers of one training run (batch rows of 256 per activation)
struct AgxTrainBuffers {
    int batch;
    long nbatches;            // training samples / batch
    const float* feat;        // [samples][4]
    const float* target;      // [samples]
    float* theta;             // parameters (fp32 master copy)
    float* grad;              // parameter gradient (same layout)
    float* m1;                // Adam moments
    float* m2;
    void* Wb;                 // bf16 W2, W3, W4 row-major (3 x 65536)
    void* H[4];               // bf16 activations h1..h4 [batch][256]
    float* S[4];              // silu' (bf16-rounded) [batch][256]
    float* Z;                 // GEMM output [batch][256]
    float* G;                 // gradient w.r.t. pre-activations [batch][256]
    void* Gb;                 // its bf16 copy
    float* dout;              // d loss / d out per sample [batch]
    long* step;               // step counter (device)
    float* lossHist;          // per-step mean squared error (device, accumulated)
};

hipError_t agxLaunchSurrogateEval(const AgxMlpDev* dM, const AgxMultipoleDev* dBase, long n,
    const double* pos, double* phi, double* grad, hipStream_t stream);
hipError_t agxLaunchMlpFeatures(const AgxMlpDev* dM, long n, const float* feat, float* out, float* dfeat,
    hipStream_t stream);
hipError_t agxLaunchGemm(int M, int N, int K, int splits, bool aK, bool bK, const void* A, long lda,
    const void* B, long ldb, float* C, long ldc, int mode, hipStream_t stream);
// one training step; with hp.lr == 0 only the loss and gradient are computed (no update)
hipError_t agxEnqueueTrainStep(const AgxTrainBuffers& T, const AgxTrainHyper& hp, hipStream_t s);
hipError_t agxLaunchPackTrain(const float* theta, void* Wb, hipStream_t s);
hipError_t agxLaunchSamplePoints(long n, unsigned long long seed, double rmin, double rmax, double* pos,
    hipStream_t s);
hipError_t agxLaunchResidual(long n, const double* pos, const double* phi, const double* phib, float lnScale,
    float* feat, double* q, double* sumsq, hipStream_t s);
hipError_t agxLaunchTarget(long n, const double* q, double inv, float* target, hipStream_t s);

// surrogate orbit integration: per-orbit fallback to the FP64 Multipole
struct AgxSurrogateOrbitArgs {
    double maxRelDE;          // fallback when |E/E0 - 1| (energy from the FP64 Multipole) exceeds this
    double rmin, rmax;        // ... or when an evaluation point leaves [rmin, rmax]
    int* fallback;            // orbit indices handed over to the FP64 kernel
    unsigned int* nfallback;  // their count (zero on entry)
    unsigned int abortAt = 0; // > 0: stop the whole launch once nfallback reaches it (pilot runs)
};
int agxOrbitSurrogateBlocksPerCU();
hipError_t agxLaunchOrbitSurrogate(const AgxMultipoleDev* dP, const AgxMultipoleDev* dBase, const AgxMlpDev* dM,
    const AgxOrbitArgs& A, const AgxSurrogateOrbitArgs& SA, int numBlocks, hipStream_t stream);
hipError_t agxLaunchMarkFallback(const int* list, long n, int* status, hipStream_t stream);
