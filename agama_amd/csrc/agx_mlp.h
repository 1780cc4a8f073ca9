// MLP potential surrogate on gfx950 MFMA: the network, its packed weight layout and the fused
// per-wavefront forward + input-gradient backward used by the evaluation and orbit kernels.
//
// Network (no counterpart in the reference; it approximates the Multipole potential of
// src/potential_multipole.cpp:1969-2080):
//   f = (|x|/r, |y|/r, |z|/r, ln r / lnScale)                                  (float)
//   h1 = silu(W1 f + b1)                          4 -> 256, float VALU
//   h_l = silu(W_l h_{l-1} + b_l), l = 2..4       256 -> 256, bf16 MFMA, fp32 accumulation
//   out = w5 . h4 + b5                            float
//   Phi = Phi_base (1 + outScale * out),  Phi_base = the FP64 Multipole truncated to l <= 2
// Activations and weights entering an MFMA are bfloat16 (round to nearest even); SiLU
// derivatives for the backward are stored as bfloat16.
//
// Register layout.  A wavefront evaluates 32 points at a time in the transposed form
// Z^T = W H^T: the MFMA A operand is a 32x16 tile of W (rows = output neurons), the B operand
// is 16 activations of each of the 32 points, and the 32x32 fp32 result tile leaves lane l
// (point n = l & 31, half g = l >> 5) holding, in register j = 0..15, output neuron
//     o(mt, g, j) = 32 mt + 8 (j >> 2) + 4 g + (j & 3)        of m-tile mt = 0..7.
// Registers 8q..8q+7 of m-tile mt are then exactly the 8 values that lane must supply as B
// operand for k-block kb = 2 mt + q of the next layer, once the next layer's weight columns
// are permuted accordingly: logical k = 16 kb + 8 g + p  <->  neuron 16 kb + 8 (p >> 2) + 4 g +
// (p & 3).  No data moves between lanes from one layer to the next.
#pragma once
#include <hip/hip_runtime.h>

#define AGX_MLP_WIDTH 256
#define AGX_MLP_NIN 4
#define AGX_MLP_NPARAM (256 * 4 + 256 + 3 * (256 * 256 + 256) + 256 + 1)

typedef __bf16 agx_bf16x8 __attribute__((ext_vector_type(8)));
typedef float agx_f32x16 __attribute__((ext_vector_type(16)));

// offsets of the tensors in the flat float parameter vector (weights row-major [out][in])
enum : int {
    AGX_P_W1 = 0, AGX_P_B1 = AGX_P_W1 + 256 * 4, AGX_P_W2 = AGX_P_B1 + 256,
    AGX_P_B2 = AGX_P_W2 + 65536, AGX_P_W3 = AGX_P_B2 + 256, AGX_P_B3 = AGX_P_W3 + 65536,
    AGX_P_W4 = AGX_P_B3 + 256, AGX_P_B4 = AGX_P_W4 + 65536, AGX_P_W5 = AGX_P_B4 + 256,
    AGX_P_B5 = AGX_P_W5 + 256, AGX_P_END = AGX_P_B5 + 1
};
static_assert(AGX_P_END == AGX_MLP_NPARAM, "parameter layout");

// neuron held by (m-tile, half, register) in the MFMA result layout
__host__ __device__ inline int agxMlpNeuron(int mt, int g, int j) { return 32 * mt + 8 * (j >> 2) + 4 * g + (j & 3); }
// neuron of logical k index within the permuted contraction
__host__ __device__ inline int agxMlpKNeuron(int k) { return (k & ~15) + 8 * ((k & 7) >> 2) + 4 * ((k >> 3) & 1) + (k & 3); }

// device-side network (all arrays in device memory; the struct itself too)
struct AgxMlpDev {
    const agx_bf16x8* fwdA[3];   // layers 2..4: [mt 8][kb 16][lane 64] tiles of W
    const agx_bf16x8* bwdA[3];   // layers 2..4: the same for W^T
    const float* w1p;            // [mt][g][j][4]
    const float* b1p;            // [mt][g][j]
    const float* bp[3];          // layers 2..4 biases, [mt][g][j]
    const float* w5p;            // [mt][g][j]
    float b5;
    float lnScale, outScale;     // feature and output scaling
    double rminTrain, rmaxTrain; // radii the network was trained on
};

// LDS the fused evaluation needs per wavefront: SiLU derivatives of layers 1..3
constexpr int AGX_MLP_LDS_VEC = 3 * 8 * 2 * 64;   // agx_bf16x8 entries (48 KB)

namespace agxd {

__device__ __forceinline__ void siluBoth(float z, float& h, float& s)
{
    const float sg = 1.f / (1.f + __expf(-z));
    h = z * sg;
    s = sg * (1.f + z * (1.f - sg));
}

typedef agx_bf16x8 __attribute__((address_space(3))) agx_lds_bf16x8;
typedef const float __attribute__((address_space(1))) agx_gbl_float;
typedef unsigned agx_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ agx_gbl_float* gF(const float* p) { return (agx_gbl_float*)p; }

// Weight tiles are read with buffer loads: the tile offset is a scalar (SGPR) operand and the
// only vector operand is the lane's 16-byte offset, so no per-tile addresses occupy VGPRs.
struct WTiles {
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ __forceinline__ explicit WTiles(const agx_bf16x8* p)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc((void*)p, 0, 8 * 16 * 64 * 16, 0x00020000)) {}
    // tile (mt, kb) of this lane
    __device__ __forceinline__ agx_bf16x8 tile(int mt, int kb, int lane) const
    {
        return __builtin_bit_cast(agx_bf16x8,
            __builtin_amdgcn_raw_buffer_load_b128(rsrc, lane * 16, (mt * 16 + kb) * 1024, 0));
    }
};

// One 256x256 product for the wavefront's 32 points: for every m-tile mt, acc = init(mt) +
// W(mt rows) x in, then epi(mt, acc).  MODE 0 issues one m-tile's 16 weight tiles and waits for
// them (fewest registers; used inside the orbit kernel).  MODE 1 works on two m-tiles at a time
// (two independent accumulation chains) and loads the next pair's tiles while the current pair
// is multiplied (about 256 more registers; used by the evaluation kernel).
template<int MODE, class Init, class Epi>
__device__ __forceinline__ void mlpProduct(const WTiles& Aw, const agx_bf16x8 (&in)[16], int lane,
    Init init, Epi epi)
{
    if constexpr(MODE == 0) {
#pragma unroll
        for(int mt = 0; mt < 8; mt++) {
            agx_f32x16 acc = init(mt);
#pragma unroll
            for(int kb = 0; kb < 16; kb++)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aw.tile(mt, kb, lane), in[kb], acc, 0, 0, 0);
            epi(mt, acc);
            // keep the next m-tile's weight loads from being hoisted here (register pressure)
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
        agx_bf16x8 P0[16], P1[16], Q0[16], Q1[16];
#pragma unroll
        for(int kb = 0; kb < 16; kb++) { P0[kb] = Aw.tile(0, kb, lane); P1[kb] = Aw.tile(1, kb, lane); }
#pragma unroll
        for(int mp = 0; mp < 8; mp += 4) {
#pragma unroll
            for(int kb = 0; kb < 16; kb++) { Q0[kb] = Aw.tile(mp + 2, kb, lane); Q1[kb] = Aw.tile(mp + 3, kb, lane); }
            {
                agx_f32x16 a0 = init(mp), a1 = init(mp + 1);
#pragma unroll
                for(int kb = 0; kb < 16; kb++) {
                    a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P0[kb], in[kb], a0, 0, 0, 0);
                    a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P1[kb], in[kb], a1, 0, 0, 0);
                }
                epi(mp, a0);
                epi(mp + 1, a1);
            }
            __builtin_amdgcn_sched_barrier(0);
            if(mp + 4 < 8) {
#pragma unroll
                for(int kb = 0; kb < 16; kb++) { P0[kb] = Aw.tile(mp + 4, kb, lane); P1[kb] = Aw.tile(mp + 5, kb, lane); }
            }
            {
                agx_f32x16 a0 = init(mp + 2), a1 = init(mp + 3);
#pragma unroll
                for(int kb = 0; kb < 16; kb++) {
                    a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Q0[kb], in[kb], a0, 0, 0, 0);
                    a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Q1[kb], in[kb], a1, 0, 0, 0);
                }
                epi(mp + 2, a0);
                epi(mp + 3, a1);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

// Fused forward + input-gradient backward for the 32 points of this wavefront: lane l works
// on point l & 31 with features feat[4] (both halves of the wavefront pass the same values).
// Returns the network output and d out / d feat in every lane (both halves agree).
// sbuf: AGX_MLP_LDS_VEC entries of LDS private to this wavefront.  MODE: see mlpProduct.
template<int MODE = 0>
__device__ __forceinline__ void mlpWave(const AgxMlpDev& M, const float feat[4],
    agx_lds_bf16x8* sbuf, float& out, float dfeat[4])
{
    const int lane = threadIdx.x & 63, g = lane >> 5;
    agx_gbl_float* w1p = gF(M.w1p);
    agx_gbl_float* b1p = gF(M.b1p);
    agx_gbl_float* w5p = gF(M.w5p);
    agx_bf16x8 ha[16], hb[16];
    // ---- layer 1 (K = 4): float VALU directly in the result layout
#pragma unroll
    for(int mt = 0; mt < 8; mt++)
#pragma unroll
        for(int q = 0; q < 2; q++) {
            agx_bf16x8 hv, sv;
#pragma unroll
            for(int p = 0; p < 8; p++) {
                const int idx = (mt * 2 + g) * 16 + 8 * q + p;
                const float z = b1p[idx] + w1p[idx * 4] * feat[0] + w1p[idx * 4 + 1] * feat[1] +
                    w1p[idx * 4 + 2] * feat[2] + w1p[idx * 4 + 3] * feat[3];
                float h, s;
                siluBoth(z, h, s);
                hv[p] = (__bf16)h;
                sv[p] = (__bf16)s;
            }
            ha[2 * mt + q] = hv;
            sbuf[(mt * 2 + q) * 64 + lane] = sv;
            __builtin_amdgcn_sched_barrier(0);
        }
    // ---- hidden layers 2 and 3: bias init, SiLU epilogue, derivatives to LDS
    auto hidden = [&](int li, const agx_bf16x8 (&in)[16], agx_bf16x8 (&outv)[16]) {
        agx_gbl_float* bias = gF(M.bp[li]);
        agx_lds_bf16x8* sslot = sbuf + (li + 1) * 1024;
        mlpProduct<MODE>(WTiles(M.fwdA[li]), in, lane,
            [&](int mt) {
                agx_f32x16 acc;
#pragma unroll
                for(int j = 0; j < 16; j++) acc[j] = bias[(mt * 2 + g) * 16 + j];
                return acc;
            },
            [&](int mt, const agx_f32x16& acc) {
#pragma unroll
                for(int q = 0; q < 2; q++) {
                    agx_bf16x8 hv, sv;
#pragma unroll
                    for(int p = 0; p < 8; p++) {
                        float h, s;
                        siluBoth(acc[8 * q + p], h, s);
                        hv[p] = (__bf16)h;
                        sv[p] = (__bf16)s;
                    }
                    outv[2 * mt + q] = hv;
                    sslot[(mt * 2 + q) * 64 + lane] = sv;
                }
            });
    };
    hidden(0, ha, hb);
    hidden(1, hb, ha);
    // ---- layer 4 and the output layer; the backward starts here: g4 = w5 * silu'(z4)
    float op = 0.f;
    {
        agx_gbl_float* b4p = gF(M.bp[2]);
        mlpProduct<MODE>(WTiles(M.fwdA[2]), ha, lane,
            [&](int mt) {
                agx_f32x16 acc;
#pragma unroll
                for(int j = 0; j < 16; j++) acc[j] = b4p[(mt * 2 + g) * 16 + j];
                return acc;
            },
            [&](int mt, const agx_f32x16& acc) {
#pragma unroll
                for(int q = 0; q < 2; q++) {
                    agx_bf16x8 gv;
#pragma unroll
                    for(int p = 0; p < 8; p++) {
                        const int idx = (mt * 2 + g) * 16 + 8 * q + p;
                        float h, s;
                        siluBoth(acc[8 * q + p], h, s);
                        const float w5 = w5p[idx];
                        op += w5 * (float)(__bf16)h;
                        gv[p] = (__bf16)(w5 * (float)(__bf16)s);
                    }
                    hb[2 * mt + q] = gv;
                }
            });
    }
    // ---- backward through layers 4 and 3 (bf16 gradients), then 2 (float, into W1^T)
    auto zero = [](int) { agx_f32x16 acc; for(int j = 0; j < 16; j++) acc[j] = 0.f; return acc; };
    auto back = [&](int li, const agx_bf16x8 (&in)[16], agx_bf16x8 (&outv)[16]) {
        const agx_lds_bf16x8* sslot = sbuf + li * 1024;
        mlpProduct<MODE>(WTiles(M.bwdA[li]), in, lane, zero,
            [&](int mt, const agx_f32x16& acc) {
#pragma unroll
                for(int q = 0; q < 2; q++) {
                    const agx_bf16x8 sv = sslot[(mt * 2 + q) * 64 + lane];
                    agx_bf16x8 gv;
#pragma unroll
                    for(int p = 0; p < 8; p++) gv[p] = (__bf16)(acc[8 * q + p] * (float)sv[p]);
                    outv[2 * mt + q] = gv;
                }
            });
    };
    back(2, hb, ha);
    back(1, ha, hb);
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f;
    mlpProduct<MODE>(WTiles(M.bwdA[0]), hb, lane, zero,
        [&](int mt, const agx_f32x16& acc) {
#pragma unroll
            for(int q = 0; q < 2; q++) {
                const agx_bf16x8 sv = sbuf[(mt * 2 + q) * 64 + lane];
#pragma unroll
                for(int p = 0; p < 8; p++) {
                    const int idx = (mt * 2 + g) * 16 + 8 * q + p;
                    const float gval = acc[8 * q + p] * (float)sv[p];
                    d0 += w1p[idx * 4] * gval;
                    d1 += w1p[idx * 4 + 1] * gval;
                    d2 += w1p[idx * 4 + 2] * gval;
                    d3 += w1p[idx * 4 + 3] * gval;
                }
            }
        });
    // the two halves of the wavefront hold complementary neurons of the same point
    out = op + __shfl_xor(op, 32) + M.b5;
    dfeat[0] = d0 + __shfl_xor(d0, 32);
    dfeat[1] = d1 + __shfl_xor(d1, 32);
    dfeat[2] = d2 + __shfl_xor(d2, 32);
    dfeat[3] = d3 + __shfl_xor(d3, 32);
}

// features of a Cartesian point and the chain rule back to Cartesian gradients
__device__ __forceinline__ void mlpFeatures(const AgxMlpDev& M, double x, double y, double z, float f[4])
{
    const double r = sqrt(x * x + y * y + z * z), ri = 1 / r;
    f[0] = (float)(fabs(x) * ri);
    f[1] = (float)(fabs(y) * ri);
    f[2] = (float)(fabs(z) * ri);
    f[3] = (float)(log(r) / (double)M.lnScale);
}

// grad of out w.r.t. (x, y, z) from d out / d f
__device__ __forceinline__ void mlpChain(const AgxMlpDev& M, double x, double y, double z,
    const float df[4], double& gx, double& gy, double& gz)
{
    const double r2 = x * x + y * y + z * z, r = sqrt(r2), ri = 1 / r, ri3 = ri / r2;
    const double ax = fabs(x), ay = fabs(y), az = fabs(z);
    const double dfx = df[0], dfy = df[1], dfz = df[2], dfl = (double)df[3] / ((double)M.lnScale * r2);
    // d(|x_i|/r)/dx_c = sign(x_i) delta_ic / r - |x_i| x_c / r^3
    const double common = (dfx * ax + dfy * ay + dfz * az) * ri3;
    gx = dfx * (x > 0 ? ri : x < 0 ? -ri : 0.) - common * x + dfl * x;
    gy = dfy * (y > 0 ? ri : y < 0 ? -ri : 0.) - common * y + dfl * y;
    gz = dfz * (z > 0 ? ri : z < 0 ? -ri : 0.) - common * z + dfl * z;
}

}  // namespace agxd
