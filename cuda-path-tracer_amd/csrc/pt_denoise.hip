// pt_denoise.hip -- learned denoiser for the progressive render: a 5-level
// convolutional encoder-decoder with skip connections over radiance, albedo
// and normal buffers, trained and run with hand-written BF16 MFMA kernels
// (FP32 accumulation) on gfx950.  No counterpart in the reference.
//
// Tensors are NHWC, bf16, channels padded to a multiple of 32 (padding is
// zero and stays zero: padded weights and biases are zero).  Every 3x3
// convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16 with
// M = output channels, N = pixels, K = (tap, input channel):
//   forward        Y = act(b + W * X), X optionally the concatenation of a
//                  2x nearest-upsampled tensor and a skip tensor;
//   backward-data  dX = W^T (*) dY, the same kernel with the taps flipped and
//                  the weight matrix transposed;
//   backward-weights dW = X^T dY, reduced over pixels through LDS tiles
//                  transposed so that K (pixels) is contiguous per lane.
// Lane layout of the MFMA (checked on the device): lane l supplies
// A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16], j<8, and receives
// C[4(l/16)+v][l%16], v<4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_denoise.h"
#include "pt_device.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- bf16
// Round to nearest even (NaN kept quiet), and back.
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    uint32_t u = uint32_t(h) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ inline bf16x8 load8(const uint16_t *p) {
    u16x8 v = *reinterpret_cast<const u16x8 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}
__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0}); }

// ---------------------------------------------------------------- convolution input
// The virtual input of a convolution at output resolution (n, h, w): padded
// channels [0, cpa) from `a` (at half resolution when a_up), then [cpa,
// cpa + cpb) from `b`.  Zero outside the image.
struct conv_src {
    const uint16_t *a;
    const uint16_t *b;
    int cpa, cpb, a_up;
    int n, h, w;
};

__device__ inline bf16x8 gather8(const conv_src& s, int n, int y, int x, int ch) {
    if(y < 0 || y >= s.h || x < 0 || x >= s.w) return zero8();
    if(ch < s.cpa) {
        size_t idx = s.a_up ? (size_t(n) * size_t(s.h >> 1) + size_t(y >> 1)) * size_t(s.w >> 1) + size_t(x >> 1)
                            : (size_t(n) * size_t(s.h) + size_t(y)) * size_t(s.w) + size_t(x);
        return load8(s.a + idx * size_t(s.cpa) + size_t(ch));
    }
    size_t idx = (size_t(n) * size_t(s.h) + size_t(y)) * size_t(s.w) + size_t(x);
    return load8(s.b + idx * size_t(s.cpb) + size_t(ch - s.cpa));
}

// The 32-channel input chunk kc of image img: its first channel at pixel 0,
// its channel stride, whether it is read at half resolution.  Offsets within
// an image fit 32 bits (the launcher checks it).
struct chunk_src { const uint16_t *p; int cp, up; };
__device__ inline chunk_src chunk_of(const conv_src& s, int img, int kc) {
    const int ch = kc * 32;
    if(ch < s.cpa) {
        const size_t hh = size_t(s.a_up ? s.h >> 1 : s.h), ww = size_t(s.a_up ? s.w >> 1 : s.w);
        return chunk_src{s.a + size_t(img) * hh * ww * size_t(s.cpa) + size_t(ch), s.cpa, s.a_up};
    }
    return chunk_src{s.b + size_t(img) * size_t(s.h) * size_t(s.w) * size_t(s.cpb) + size_t(ch - s.cpa), s.cpb, 0};
}

constexpr float LOG_CLAMP = 12.0f;  // log radiance above which the output is clamped
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// Two floats to packed bf16, rounded to nearest even (one v_cvt_pk_bf16_f32;
// equal to to_bf16 for every non-NaN value).
__device__ inline uint32_t pack_bf16(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

// Master channel of padded input channel pi (channels of a, then of b), or -1
// for padding.
__device__ inline int real_channel(int pi, int cin_a, int cpa, int cin_b) {
    if(pi < cpa) return pi < cin_a ? pi : -1;
    return pi - cpa < cin_b ? cin_a + (pi - cpa) : -1;
}

// EPI_RADIANCE (the output layer at inference): the denoised linear radiance
// max(expm1(min(y + x, LOG_CLAMP)), 0) of channels 0..2 (y the layer's output,
// x the network input, i.e. log1p of the noisy radiance) for the pixels
// inside the unpadded frame, written as 3 floats per pixel.
enum epilogue { EPI_BF16 = 0, EPI_BF16_RELU = 1, EPI_F32 = 2, EPI_RADIANCE = 3 };

// A workgroup computes 16 x 16 pixel tiles of the images for MT 16-channel
// output tiles (blockIdx.y selects which); wk is [9][kchunks][cpo][32] bf16.
// The grid is persistent: workgroup b takes tiles b, b + gridDim.x, ... and
// walks the flattened sequence of (tile, 32-channel input chunk) items.  An
// item's 18 x 18 input halo is staged through LDS (double-buffered: the next
// item's global loads, of the same tile or of the next one, are in flight
// while the current one is multiplied), laid out [channel group of 8][halo
// pixel] so that the 16 pixels of an MFMA B fragment read 256 contiguous
// bytes.  Wave w owns tile rows 4w..4w+3 (one 16-pixel N tile per row).  A
// tile's outputs are written after its last chunk, while the next tile's
// first chunk loads.  Bias is added to the channels below cout; EPI_F32
// stores only the 4-channel groups that start below cout.
constexpr int TILE = 16, HALO = TILE + 2, HALO_PIX = HALO * HALO;
constexpr int STAGE_ITEMS = HALO_PIX * 4;                       // 16-byte items per chunk
constexpr int STAGE_PER_THREAD = (STAGE_ITEMS + 255) / 256;

// Resident workgroups per compute unit: three (the most the LDS allows, at
// most 168 VGPRs and AGPRs together per lane) for up to two output tiles,
// two for more (their accumulators need the registers).
template<int MT>
constexpr int conv_waves = MT <= 2 ? 3 : 2;

template<int MT, int EPI>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(conv_waves<MT>))) void conv_tile_kernel(
        conv_src src, const uint16_t *__restrict__ wk, int kchunks, int cpo, const float *__restrict__ bias, int cout,
        void *__restrict__ out, int tiles_x, int tiles_y, const uint16_t *__restrict__ xin, int frame_h, int frame_w,
        uint16_t *__restrict__ pooled) {
    __shared__ u16x8 halo[2][4][HALO_PIX];
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int kq = lane >> 4, col = lane & 15;
    const int tiles = src.n * tiles_x * tiles_y;
    const int stride = int(gridDim.x);
    const int co0 = int(blockIdx.y) * MT * 16;
    struct place { int img, x0, y0; };
    auto where = [&](int t) {
        place w;
        w.img = t / (tiles_x * tiles_y);
        const int r = t - w.img * tiles_x * tiles_y;
        w.y0 = (r / tiles_x) * TILE;
        w.x0 = (r - (r / tiles_x) * tiles_x) * TILE;
        return w;
    };

    u16x8 staged[STAGE_PER_THREAD];
    auto fetch = [&](const place& w, int kc) {
        const chunk_src c = chunk_of(src, w.img, kc);
        const int row = c.up ? src.w >> 1 : src.w;
        // a tile whose halo lies inside the image needs no bounds tests
        const bool inside = w.x0 > 0 && w.y0 > 0 && w.x0 + TILE < src.w && w.y0 + TILE < src.h;
        for(int i = 0; i < STAGE_PER_THREAD; i++) {
            const int item = tid + 256 * i;
            if(item < STAGE_ITEMS) {
                const int hp = item >> 2, q = item & 3;
                const int hy = hp / HALO, hx = hp - hy * HALO;
                const int y = w.y0 - 1 + hy, x = w.x0 - 1 + hx;
                const bool ok = inside || (y >= 0 && y < src.h && x >= 0 && x < src.w);
                const int pix = c.up ? (y >> 1) * row + (x >> 1) : y * row + x;
                staged[i] = ok ? *reinterpret_cast<const u16x8 *>(c.p + pix * c.cp + 8 * q) : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
        }
    };
    auto store = [&](int buf) {
        for(int i = 0; i < STAGE_PER_THREAD; i++) {
            const int item = tid + 256 * i;
            if(item < STAGE_ITEMS) halo[buf][item & 3][item >> 2] = staged[i];
        }
    };

    f32x4 acc[MT][4];
    for(int mt = 0; mt < MT; mt++)
        for(int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0, 0, 0, 0};
    // this lane's four output channels per output tile, and their biases
    float bv[MT][4];
    for(int mt = 0; mt < MT; mt++)
        for(int j = 0; j < 4; j++) {
            const int co = co0 + mt * 16 + 4 * kq + j;
            bv[mt][j] = (bias && co < cout) ? bias[co] : 0.0f;
        }

    int tile = int(blockIdx.x), kc = 0;
    if(tile >= tiles) return;  // whole workgroup
    place cur = where(tile);
    fetch(cur, 0);
    store(0);
    __syncthreads();
    for(int buf = 0;; buf ^= 1) {
        // the item after this one: the next chunk, or the next tile's first
        int next_tile = tile, next_kc = kc + 1;
        if(next_kc == kchunks) { next_kc = 0; next_tile = tile + stride; }
        const bool more = next_tile < tiles;
        const place nxt = more ? where(next_tile) : cur;
        if(more) fetch(nxt, next_kc);
        const uint16_t *wchunk = wk + (size_t(kc) * size_t(cpo) + size_t(co0 + col)) * 32 + size_t(8 * kq);
#pragma unroll
        for(int t = 0; t < 9; t++) {
            const int dy = t / 3, dx = t % 3;
            bf16x8 afrag[MT];
            const uint16_t *wrow = wchunk + size_t(t) * size_t(kchunks) * size_t(cpo) * 32;
            for(int mt = 0; mt < MT; mt++) afrag[mt] = load8(wrow + size_t(mt) * 16 * 32);
            bf16x8 bfrag[4];
            for(int nt = 0; nt < 4; nt++)
                bfrag[nt] = __builtin_bit_cast(bf16x8, halo[buf][kq][(4 * wave + nt + dy) * HALO + col + dx]);
            for(int mt = 0; mt < MT; mt++)
                for(int nt = 0; nt < 4; nt++)
                    acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag[mt], bfrag[nt], acc[mt][nt], 0, 0, 0);
        }
        if(kc == kchunks - 1) {  // the tile is complete: bias, activation, store
            const int x = cur.x0 + col;
            uint32_t packed[MT][4][2];  // bf16 outputs (EPI_BF16*), kept for the pooled output
            for(int nt = 0; nt < 4; nt++) {
                const int y = cur.y0 + 4 * wave + nt;
                const bool inside = x < src.w && y < src.h;
                const size_t p = (size_t(cur.img) * size_t(src.h) + size_t(y)) * size_t(src.w) + size_t(x);
                for(int mt = 0; mt < MT; mt++) {
                    const int co = co0 + mt * 16 + 4 * kq;
                    float v[4];
                    for(int j = 0; j < 4; j++) {
                        v[j] = acc[mt][nt][j];
                        if(bias) v[j] += bv[mt][j];
                        if(EPI == EPI_BF16_RELU) v[j] = fmaxf(v[j], 0.0f);
                    }
                    if(EPI == EPI_RADIANCE) {
                        // channels 0..2, held by the first 4-channel group (predicted_log)
                        if(co == 0 && x < frame_w && y < frame_h) {
                            float *rad = static_cast<float *>(out) +
                                         ((size_t(cur.img) * size_t(frame_h) + size_t(y)) * size_t(frame_w) + size_t(x)) * 3;
                            for(int c = 0; c < 3; c++) {
                                const float pred = v[c] + from_bf16(xin[p * 32 + size_t(c)]);
                                rad[c] = fmaxf(expm1f(fminf(pred, LOG_CLAMP)), 0.0f);
                            }
                        }
                    } else if(EPI == EPI_F32) {
                        if(co < cout && inside)
                            *reinterpret_cast<float4 *>(static_cast<float *>(out) + p * size_t(cpo) + size_t(co)) =
                                make_float4(v[0], v[1], v[2], v[3]);
                    } else {
                        packed[mt][nt][0] = pack_bf16(v[0], v[1]);
                        packed[mt][nt][1] = pack_bf16(v[2], v[3]);
                        if(inside)
                            *reinterpret_cast<uint2 *>(static_cast<uint16_t *>(out) + p * size_t(cpo) + size_t(co)) =
                                make_uint2(packed[mt][nt][0], packed[mt][nt][1]);
                    }
                }
            }
            if constexpr(EPI == EPI_BF16 || EPI == EPI_BF16_RELU) {
                // The 2x2 average pooling of the stored (bf16) values into the
                // half-resolution tensor `pooled` (floor of the size): rows np,
                // np + 1 of this lane and columns col, col ^ 1 (the neighbouring
                // lane), summed left to right then top to bottom, times 0.25,
                // rounded to bf16.
                if(pooled) {
                    const int ho = src.h >> 1, wo = src.w >> 1;
                    const int X = x >> 1;
                    for(int np = 0; np < 4; np += 2) {
                        const int Y = (cur.y0 + 4 * wave + np) >> 1;
                        for(int mt = 0; mt < MT; mt++) {
                            uint32_t res[2];
                            for(int h2 = 0; h2 < 2; h2++) {
                                const uint32_t t0 = packed[mt][np][h2], t1 = packed[mt][np + 1][h2];
                                const uint32_t u0 = uint32_t(__shfl_xor(int(t0), 1)), u1 = uint32_t(__shfl_xor(int(t1), 1));
                                float avg[2];
                                for(int e = 0; e < 2; e++) {
                                    const int sh = 16 * e;
                                    const float a00 = from_bf16(uint16_t(t0 >> sh)), a01 = from_bf16(uint16_t(u0 >> sh));
                                    const float a10 = from_bf16(uint16_t(t1 >> sh)), a11 = from_bf16(uint16_t(u1 >> sh));
                                    avg[e] = 0.25f * (a00 + a01 + a10 + a11);
                                }
                                res[h2] = pack_bf16(avg[0], avg[1]);
                            }
                            if(!(col & 1) && X < wo && Y < ho) {
                                const size_t q = (size_t(cur.img) * size_t(ho) + size_t(Y)) * size_t(wo) + size_t(X);
                                *reinterpret_cast<uint2 *>(pooled + q * size_t(cpo) + size_t(co0 + mt * 16 + 4 * kq)) =
                                    make_uint2(res[0], res[1]);
                            }
                        }
                    }
                }
            }
            for(int mt = 0; mt < MT; mt++)
                for(int nt = 0; nt < 4; nt++) acc[mt][nt] = f32x4{0, 0, 0, 0};
        }
        if(!more) break;
        store(buf ^ 1);
        __syncthreads();
        tile = next_tile; kc = next_kc; cur = nxt;
    }
}

// dW partials: for input chunk kc (blockIdx.y) and pixel split blockIdx.x,
// partial[s][t][kc*32 + ci][co] = sum over the split's pixels p of
// X(p + offset(t))[ci] * dY[p][co], all nine taps t at once so that each
// 64-pixel sub-block of dY is staged once.  Both operands are transposed
// through LDS so that K (pixels) is contiguous per lane.  NCO = cpo / 16
// output-channel tiles; the 18 (tap, 16-channel input half) pairs are dealt
// to the 4 waves, each wave computing all NCO output tiles of its pairs: the
// dY fragments stay in registers for the whole sub-block and each X fragment
// serves NCO MFMAs, so that the LDS reads per MFMA fall from 2 to about
// 0.5-0.7.  Every tile accumulates its MFMAs in the same order (sub-block,
// then the two 32-pixel steps).
// The chunk-0 blocks also sum dY over the split's pixels into
// bpartial[s][co] (the bias gradient's partials): each thread accumulates
// the 8-channel groups it stages, in pixel order, and the threads' sums of a
// channel are added in thread order at the end.
constexpr int WG_PIX = 64;
constexpr int LDS_ROW = WG_PIX + 8;  // bf16 per LDS row (16-byte aligned rows, fewer bank conflicts)

// Column swizzles of the transposed tiles.  A row holds one channel's 64
// pixels; 8 rows of 36 dwords wrap the LDS banks exactly, so the 8-channel
// groups a wave writes at once (one row each, the same pixels) would land
// on the same banks.  Each row's pixel column is XORed with a multiple of 8
// pixels that depends on its channel group, which spreads those writes over
// the banks and keeps every aligned 8-pixel block (an MFMA operand read)
// contiguous.  X tile: a wave writes 16 pixels of 4 groups; dY tile: 64 / G
// pixels of G = cpo / 8 groups.
__device__ inline int x_swizzle(int row) { return 16 * ((row >> 3) & 3); }
template<int G> __device__ inline int dy_swizzle(int row) {
    return G == 4 ? 16 * ((row >> 3) & 3) : 8 * ((row >> 3) & 7);
}

template<int NCO>
__global__ __launch_bounds__(256) void wgrad_kernel(conv_src src, const uint16_t *__restrict__ dy, int cpo,
                                                    int64_t blocks_per_split, int cpi, float *__restrict__ partial,
                                                    float *__restrict__ bpartial) {
    __shared__ __attribute__((aligned(16))) uint16_t xt[9 * 32 * LDS_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t dt[NCO * 16 * LDS_ROW];
    const int tid = int(threadIdx.x), lane = tid & 63, wave = tid >> 6;
    const int kq = lane >> 4, col = lane & 15;
    const int kc = int(blockIdx.y);
    const int64_t total = int64_t(src.n) * src.h * src.w;
    const int64_t hw = int64_t(src.h) * src.w;
    const int64_t nsub = (total + WG_PIX - 1) / WG_PIX;
    const int64_t sb0 = int64_t(blockIdx.x) * blocks_per_split;
    const int64_t sb1 = std::min(nsub, sb0 + blocks_per_split);
    constexpr int PAIRS = 9 * 2;
    constexpr int PPW = (PAIRS + 3) / 4;
    f32x4 acc[PPW][NCO];
    for(int i = 0; i < PPW; i++)
        for(int tj = 0; tj < NCO; tj++) acc[i][tj] = f32x4{0, 0, 0, 0};
    const int groups = cpo / 8;
    const int pxl = tid >> 2, g = tid & 3;  // this thread's pixel and 8-channel group of the X tile
    // bias sums of the dY groups this thread stages: item tid + 256 j has group (tid + 256 j) % groups
    constexpr int BJ = (WG_PIX * 2 * NCO + 255) / 256;
    float bsum[BJ][8];
    for(int j = 0; j < BJ; j++)
        for(int e = 0; e < 8; e++) bsum[j][e] = 0.0f;
    for(int64_t sb = sb0; sb < sb1; sb++) {
        const int64_t p0 = sb * WG_PIX;
        {   // X at the nine tap offsets, transposed into xt[t][ci][px].
            const int64_t p = p0 + pxl;
            int n = 0, y = 0, x = 0;
            if(p < total) {
                n = int(p / hw);
                const int64_t r = p - int64_t(n) * hw;
                y = int(r / src.w);
                x = int(r - int64_t(y) * src.w);
            }
            for(int t = 0; t < 9; t++) {
                bf16x8 v = zero8();
                if(p < total) v = gather8(src, n, y + t / 3 - 1, x + t % 3 - 1, kc * 32 + 8 * g);
                const u16x8 u = __builtin_bit_cast(u16x8, v);
                for(int j = 0; j < 8; j++) xt[(t * 32 + 8 * g + j) * LDS_ROW + (pxl ^ (16 * g))] = u[j];
            }
        }
        for(int j = 0; j < BJ; j++) {  // dY tile, transposed into dt[co][px]
            const int i = tid + 256 * j;
            if(i >= WG_PIX * groups) break;
            const int q = i / groups, gg = i - q * groups;
            const int64_t p = p0 + q;
            u16x8 u = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if(p < total) u = *reinterpret_cast<const u16x8 *>(dy + size_t(p) * size_t(cpo) + size_t(8 * gg));
            for(int e = 0; e < 8; e++) dt[(8 * gg + e) * LDS_ROW + (q ^ dy_swizzle<2 * NCO>(8 * gg))] = u[e];
            if(kc == 0)
                for(int e = 0; e < 8; e++) bsum[j][e] += from_bf16(u[e]);
        }
        __syncthreads();
        constexpr int KS = WG_PIX / 32;
        bf16x8 b[NCO][KS];
        for(int tj = 0; tj < NCO; tj++)
            for(int ks = 0; ks < KS; ks++)
                b[tj][ks] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(
                    &dt[(tj * 16 + col) * LDS_ROW + ((ks * 32 + 8 * kq) ^ dy_swizzle<2 * NCO>(tj * 16 + col))]));
        for(int i = 0; i < PPW; i++) {
            const int pair = wave + 4 * i;
            if(pair >= PAIRS) break;
            const int t = pair >> 1, ti = pair & 1;
            for(int ks = 0; ks < KS; ks++) {
                const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(
                    &xt[(t * 32 + ti * 16 + col) * LDS_ROW + ((ks * 32 + 8 * kq) ^ x_swizzle(ti * 16 + col))]));
                for(int tj = 0; tj < NCO; tj++)
                    acc[i][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[tj][ks], acc[i][tj], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    for(int i = 0; i < PPW; i++) {
        const int pair = wave + 4 * i;
        if(pair >= PAIRS) break;
        const int t = pair >> 1, ti = pair & 1;
        for(int tj = 0; tj < NCO; tj++)
            for(int v = 0; v < 4; v++) {
                const int ci = kc * 32 + ti * 16 + 4 * kq + v, co = tj * 16 + col;
                partial[((size_t(blockIdx.x) * 9 + size_t(t)) * size_t(cpi) + size_t(ci)) * size_t(cpo) + size_t(co)] =
                    acc[i][tj][v];
            }
    }
    if(kc == 0) {
        // the threads' bias sums through LDS (xt is free after the loop's last barrier)
        float *red = reinterpret_cast<float *>(xt);
        static_assert(BJ * 256 * 8 * sizeof(float) <= sizeof(xt), "bias sums fit in the X tile");
        for(int j = 0; j < BJ; j++)
            for(int e = 0; e < 8; e++) red[(j * 256 + tid) * 8 + e] = bsum[j][e];
        __syncthreads();
        if(tid < cpo) {
            // item q * groups + gg (sub-block pixel q, group gg) was staged by
            // thread (q * groups + gg) % 256, iteration (q * groups + gg) / 256
            const int gg = tid >> 3, e = tid & 7;
            float b = 0.0f;
            for(int q = 0; q < WG_PIX; q++) b += red[(q * groups + gg) * 8 + e];
            bpartial[size_t(blockIdx.x) * size_t(cpo) + size_t(tid)] = b;
        }
    }
}

// Sum the dW partials into the FP32 gradient in the master layout [co][t][ci].
// Sum the dW partials over the pixel splits into the FP32 gradient in the
// master layout [co][t][ci].  Threads run along the partial layout
// [t][pi][co] (coalesced reads); threadIdx.y takes every fourth split and the
// four sums are combined in a fixed order.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float *partial, int splits, int cpi, int cpo, int cin_a,
                                                           int cpa, int cin_b, int cout, float *grad) {
    __shared__ float part[4][64];
    const int64_t per_split = int64_t(9) * cpi * cpo;
    const int64_t j = int64_t(blockIdx.x) * 64 + threadIdx.x;
    const int sg = int(threadIdx.y);
    float s = 0;
    if(j < per_split)
        for(int k = sg; k < splits; k += 4) s += partial[size_t(k) * size_t(per_split) + size_t(j)];
    part[sg][threadIdx.x] = s;
    __syncthreads();
    if(sg != 0 || j >= per_split) return;
    const int co = int(j % cpo), pi = int((j / cpo) % cpi), t = int(j / (int64_t(cpo) * cpi));
    const int ci = real_channel(pi, cin_a, cpa, cin_b);
    if(co >= cout || ci < 0) return;
    const float total = (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
    grad[(size_t(co) * 9 + size_t(t)) * size_t(cin_a + cin_b) + size_t(ci)] = total;
}

// One block per output channel: a fixed-order tree sum of the block partials.
__global__ __launch_bounds__(256) void bgrad_reduce_kernel(const float *partial, int blocks, int cpo, int cout, float *grad) {
    __shared__ float red[256];
    const int co = int(blockIdx.x), tid = int(threadIdx.x);
    float s = 0;
    for(int k = tid; k < blocks; k += 256) s += partial[size_t(k) * size_t(cpo) + size_t(co)];
    red[tid] = s;
    __syncthreads();
    for(int w = 128; w > 0; w >>= 1) {
        if(tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if(tid == 0 && co < cout) grad[co] = red[0];
}

// ---------------------------------------------------------------- elementwise stages
// Gradient routing and ReLU masks.  Every tensor has a multiple of 32
// channels per pixel, so each thread moves one 4-channel group (16 bytes) or,
// for the bf16 masks, 8 channels; consecutive threads take consecutive groups.
// Per element the arithmetic is the plain one (g += dx, g += 0.25 * dx, the
// 2x2 sum left to right then top to bottom).  With `first` set the kernel is
// the gradient's first contribution and g is read as zero (0 + v, so that a
// -0 contribution stores +0 as an add into a zeroed buffer would): the
// buffers need no clearing.

// i / d for a 64-bit i and a 32-bit d, in 32-bit arithmetic while i fits.
__device__ inline uint64_t udiv(uint64_t i, uint32_t d) { return (i >> 32) ? i / d : uint64_t(uint32_t(i) / d); }

__device__ inline float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ inline float4 accumulate(const float4 *g, uint64_t i, bool first, float4 v) {
    return add4(first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : g[i], v);
}

// g[p][c] += dx[p][c0 + c] for c < cp (dx has dxc channels per pixel).
__global__ void route_same_kernel(float *g, const float *dx, int64_t pixels, int cp, int dxc, int c0, int first) {
    const uint32_t gp = uint32_t(cp) / 4u;
    const uint64_t count = uint64_t(pixels) * gp;
    float4 *g4 = reinterpret_cast<float4 *>(g);
    for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t p = udiv(i, gp);
        const uint32_t c = uint32_t(i - p * gp) * 4u;
        g4[i] = accumulate(g4, i, first, *reinterpret_cast<const float4 *>(dx + p * uint64_t(dxc) + uint64_t(c0) + c));
    }
}
// Gradient of a 2x nearest upsample: g[q][c] += sum of the 2x2 block of dx
// (full resolution (n, h, w), dxc channels, channels [0, cp)).
__global__ void route_up_kernel(float *g, const float *dx, int n, int h, int w, int cp, int dxc, int first) {
    const uint32_t ho = uint32_t(h) >> 1, wo = uint32_t(w) >> 1, gp = uint32_t(cp) / 4u;
    const uint64_t count = uint64_t(n) * ho * wo * gp;
    float4 *g4 = reinterpret_cast<float4 *>(g);
    for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t q = udiv(i, gp);
        const uint32_t c = uint32_t(i - q * gp) * 4u;
        const uint64_t row = udiv(q, wo);
        const uint32_t x = uint32_t(q - row * wo);
        const uint64_t b = udiv(row, ho);
        const uint32_t y = uint32_t(row - b * ho);
        const uint64_t r0 = (b * uint64_t(h) + 2u * y) * uint64_t(w) + 2u * x, r1 = r0 + uint64_t(w);
        const float4 d00 = *reinterpret_cast<const float4 *>(dx + r0 * uint64_t(dxc) + c);
        const float4 d01 = *reinterpret_cast<const float4 *>(dx + (r0 + 1) * uint64_t(dxc) + c);
        const float4 d10 = *reinterpret_cast<const float4 *>(dx + r1 * uint64_t(dxc) + c);
        const float4 d11 = *reinterpret_cast<const float4 *>(dx + (r1 + 1) * uint64_t(dxc) + c);
        g4[i] = accumulate(g4, i, first, add4(add4(add4(d00, d01), d10), d11));
    }
}
// Gradient of a 2x2 average pool: g[p][c] += dx[pool(p)][c] / 4 (g at full
// resolution (n, h, w), dx at half, both cp channels); one thread per pooled
// group, updating its four full-resolution pixels.
__global__ void route_pool_kernel(float *g, const float *dx, int n, int h, int w, int cp, int first) {
    const uint32_t ho = uint32_t(h) >> 1, wo = uint32_t(w) >> 1, gp = uint32_t(cp) / 4u;
    const uint64_t count = uint64_t(n) * ho * wo * gp;
    float4 *g4 = reinterpret_cast<float4 *>(g);
    for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
        const uint64_t q = udiv(i, gp);
        const uint32_t k = uint32_t(i - q * gp);
        const uint64_t row = udiv(q, wo);
        const uint32_t x = uint32_t(q - row * wo);
        const uint64_t b = udiv(row, ho);
        const uint32_t y = uint32_t(row - b * ho);
        const float4 d = reinterpret_cast<const float4 *>(dx)[i];
        const float4 v = make_float4(0.25f * d.x, 0.25f * d.y, 0.25f * d.z, 0.25f * d.w);
        const uint64_t p0 = (b * uint64_t(h) + 2u * y) * uint64_t(w) + 2u * x, p1 = p0 + uint64_t(w);
        const uint64_t fine[4] = {p0, p0 + 1, p1, p1 + 1};
        for(int j = 0; j < 4; j++) g4[fine[j] * gp + k] = accumulate(g4, fine[j] * gp + k, first, v);
    }
}
// dY (bf16) of the layer that produced y: g where y > 0 (ReLU), else 0;
// count a multiple of 8.
__global__ void relu_grad_kernel(uint16_t *dy, const float *g, const uint16_t *y, int64_t count) {
    const uint64_t n8 = uint64_t(count) / 8u;
    for(uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n8; i += uint64_t(gridDim.x) * blockDim.x) {
        const u16x8 yv = reinterpret_cast<const u16x8 *>(y)[i];
        const float4 g0 = reinterpret_cast<const float4 *>(g)[2 * i], g1 = reinterpret_cast<const float4 *>(g)[2 * i + 1];
        const float gv[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
        u16x8 out;
        for(int j = 0; j < 8; j++) out[j] = from_bf16(yv[j]) > 0.0f ? to_bf16(gv[j]) : uint16_t(0);
        reinterpret_cast<u16x8 *>(dy)[i] = out;
    }
}

// Network input: log1p(radiance), albedo, normal (9 channels of 32).
__device__ inline void write_input(uint16_t *x, const float f[9]) {
    u16x8 a, b, z = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    for(int c = 0; c < 3; c++) a[c] = to_bf16(log1pf(fmaxf(f[c], 0.0f)));
    for(int c = 3; c < 8; c++) a[c] = to_bf16(f[c]);
    b = z;
    b[0] = to_bf16(f[8]);
    u16x8 *o = reinterpret_cast<u16x8 *>(x);
    o[0] = a; o[1] = b; o[2] = z; o[3] = z;
}
// From host-layout features [p][9] (linear radiance, albedo, normal).
__global__ void features_kernel(const float *feat, int64_t pixels, uint16_t *x) {
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < pixels; p += int64_t(gridDim.x) * blockDim.x) {
        float f[9];
        for(int c = 0; c < 9; c++) f[c] = feat[size_t(p) * 9 + size_t(c)];
        write_input(x + size_t(p) * 32, f);
    }
}
// From a renderer's sums (radiance, albedo, normal scaled by k = 1/samples),
// into an image padded to (hp, wp) with zeros.
__global__ void frame_features_kernel(const float *sum, const float *aov, float k, int w, int h, int wp, int hp,
                                      uint16_t *x) {
    const int64_t count = int64_t(wp) * hp;
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < count; p += int64_t(gridDim.x) * blockDim.x) {
        const int xx = int(p % wp), yy = int(p / wp);
        float f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if(xx < w && yy < h) {
            const size_t i = size_t(yy) * size_t(w) + size_t(xx);
            for(int c = 0; c < 3; c++) f[c] = sum[3 * i + size_t(c)] * k;
            if(aov) for(int c = 0; c < 6; c++) f[3 + c] = aov[6 * i + size_t(c)] * k;
        }
        write_input(x + size_t(p) * 32, f);
    }
}
// Training batch: crops ids[b] of the resident data set.
__global__ void gather_kernel(const float *feat, const float *target, const int32_t *ids, int batch, int64_t hw,
                              uint16_t *x, float *t) {
    const int64_t count = int64_t(batch) * hw;
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x) {
        const int b = int(i / hw);
        const int64_t p = i - int64_t(b) * hw;
        const size_t src = size_t(ids[b]) * size_t(hw) + size_t(p);
        float f[9];
        for(int c = 0; c < 9; c++) f[c] = feat[src * 9 + size_t(c)];
        write_input(x + size_t(i) * 32, f);
        for(int c = 0; c < 3; c++) t[size_t(i) * 3 + size_t(c)] = target[src * 3 + size_t(c)];
    }
}

// The network predicts a correction to the input's log radiance.
__device__ inline float predicted_log(const float *out, const uint16_t *x, int64_t p, int c) {
    return out[size_t(p) * 32 + size_t(c)] + from_bf16(x[size_t(p) * 32 + size_t(c)]);
}
constexpr float RELMSE_EPS = 0.01f;

// Loss = mean |y - log1p(t)| + mean (expm1(y) - t)^2 / (t^2 + 0.01) over
// pixels and channels; writes dL/d(out) (bf16, 32 channels) and per-block
// loss sums.
__global__ __launch_bounds__(256) void loss_kernel(const float *out, const uint16_t *x, const float *target,
                                                   int64_t pixels, uint16_t *dout, float *block_loss) {
    __shared__ float red[256];
    const float scale = 1.0f / float(3 * pixels);
    float s = 0;
    for(int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < pixels; p += int64_t(gridDim.x) * blockDim.x) {
        u16x8 g0 = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        for(int c = 0; c < 3; c++) {
            const float y = predicted_log(out, x, p, c);
            const float t = fmaxf(target[size_t(p) * 3 + size_t(c)], 0.0f);
            const float lt = log1pf(t);
            const float yc = fminf(y, LOG_CLAMP);
            const float e = expm1f(yc);
            const float den = t * t + RELMSE_EPS;
            s += fabsf(y - lt) + (e - t) * (e - t) / den;
            float grad = (y > lt ? 1.0f : (y < lt ? -1.0f : 0.0f));
            if(y < LOG_CLAMP) grad += 2.0f * (e - t) / den * (e + 1.0f);
            g0[c] = to_bf16(grad * scale);
        }
        u16x8 *o = reinterpret_cast<u16x8 *>(dout + size_t(p) * 32);
        const u16x8 z = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
        o[0] = g0; o[1] = z; o[2] = z; o[3] = z;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for(int k = 128; k > 0; k >>= 1) {
        if(int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if(threadIdx.x == 0) block_loss[blockIdx.x] = red[0] * scale;
}
// The sum of n values in a fixed order (one block of 64 threads): thread t
// adds values t, t + 64, ... in turn, then a fixed tree over the threads.
__global__ __launch_bounds__(64) void sum_kernel(const float *v, int n, float *out) {
    __shared__ float part[64];
    float s = 0;
    for(int i = int(threadIdx.x); i < n; i += 64) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for(int k = 32; k > 0; k >>= 1) {
        if(int(threadIdx.x) < k) part[threadIdx.x] += part[threadIdx.x + k];
        __syncthreads();
    }
    if(threadIdx.x == 0) *out = part[0];
}

// Adam on the FP32 master parameters.
__global__ void adam_kernel(float *w, const float *g, float *m, float *v, int64_t count, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1 - b1) * gi;
        const float vi = b2 * v[i] + (1 - b2) * gi * gi;
        m[i] = mi; v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// Packed bf16 weights from the master layout W[co][t][ci].
// Forward: wk[t][kc][co][k] = W[co][t][ci(kc*32 + k)].
__global__ void pack_forward_kernel(const float *w, int cin_a, int cpa, int cin_b, int cpb, int cout, int cpo, uint16_t *wk) {
    const int cpi = cpa + cpb, cin = cin_a + cin_b;
    const int64_t count = int64_t(9) * cpi * cpo;
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x) {
        const int k = int(i % 32);
        const int co = int((i / 32) % cpo);
        const int kc = int((i / (32 * int64_t(cpo))) % (cpi / 32));
        const int t = int(i / (int64_t(cpo) * cpi));
        const int ci = real_channel(kc * 32 + k, cin_a, cpa, cin_b);
        const float v = (co < cout && ci >= 0) ? w[(size_t(co) * 9 + size_t(t)) * size_t(cin) + size_t(ci)] : 0.0f;
        wk[size_t(i)] = to_bf16(v);
    }
}
// Backward-data: wt[t'][oc][pi][k] = W[co = oc*32 + k][8 - t'][ci(pi)].
__global__ void pack_backward_kernel(const float *w, int cin_a, int cpa, int cin_b, int cpb, int cout, int cpo, uint16_t *wt) {
    const int cpi = cpa + cpb, cin = cin_a + cin_b;
    const int64_t count = int64_t(9) * cpo * cpi;
    for(int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x) {
        const int k = int(i % 32);
        const int pi = int((i / 32) % cpi);
        const int oc = int((i / (32 * int64_t(cpi))) % (cpo / 32));
        const int tp = int(i / (int64_t(cpi) * cpo));
        const int co = oc * 32 + k, t = 8 - tp;
        const int ci = real_channel(pi, cin_a, cpa, cin_b);
        const float v = (co < cout && ci >= 0) ? w[(size_t(co) * 9 + size_t(t)) * size_t(cin) + size_t(ci)] : 0.0f;
        wt[size_t(i)] = to_bf16(v);
    }
}

// ---------------------------------------------------------------- host side
void hip_check(hipError_t e, const char *what) {
    if(e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HC(x) hip_check((x), #x)

template<class T> struct dbuf {
    T *p = nullptr;
    size_t n = 0;
    dbuf() = default;
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    ~dbuf() { if(p) (void)hipFree(p); }
    void ensure(size_t count) {
        if(count <= n) return;
        if(p) HC(hipFree(p));
        p = nullptr; n = 0;
        HC(hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(count, 1) * sizeof(T)));
        n = count;
    }
};

int pad32(int c) { return (c + 31) / 32 * 32; }
unsigned grid_for(int64_t n, int block = 256) {
    return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 65536)));
}

// Tensors of the network: input, ten hidden outputs, four pooled tensors.
enum tensor_id { T_X = 0, T_Y0 = 1, T_P1 = 11, T_P2, T_P3, T_P4, T_COUNT };
inline int t_y(int k) { return T_Y0 + k; }

struct layer_spec {
    int src_a, src_b;   // tensor ids (src_b < 0: plain input)
    int cin_a, cin_b, cout;
    bool relu, up_a;
    int level;
};

// 5 levels, channels 32-48-64-80-96; decoder inputs concatenate the
// upsampled deeper output with the encoder output of the same level.
const layer_spec SPECS[11] = {
    {T_X, -1, 9, 0, 32, true, false, 0},
    {T_Y0, -1, 32, 0, 32, true, false, 0},
    {T_P1, -1, 32, 0, 48, true, false, 1},
    {T_P2, -1, 48, 0, 64, true, false, 2},
    {T_P3, -1, 64, 0, 80, true, false, 3},
    {T_P4, -1, 80, 0, 96, true, false, 4},
    {T_Y0 + 5, T_Y0 + 4, 96, 80, 80, true, true, 3},
    {T_Y0 + 6, T_Y0 + 3, 80, 64, 64, true, true, 2},
    {T_Y0 + 7, T_Y0 + 2, 64, 48, 48, true, true, 1},
    {T_Y0 + 8, T_Y0 + 1, 48, 32, 32, true, true, 0},
    {T_Y0 + 9, -1, 32, 0, 3, false, false, 0},
};
constexpr int NLAYERS = 11;

int64_t layer_params(const layer_spec& l) { return int64_t(l.cout) * 9 * (l.cin_a + l.cin_b) + l.cout; }

} // namespace

struct ptd_network {
    int64_t offsets[NLAYERS + 1];    // master parameter offset of each layer (W then b)
    dbuf<float> params, grads, adam_m, adam_v;
    dbuf<uint16_t> wk[NLAYERS], wt[NLAYERS];
    int cpa[NLAYERS], cpb[NLAYERS], cpo[NLAYERS];
    int64_t step = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int compute_units = 1;
    // shape-dependent buffers
    int n = 0, h = 0, w = 0;
    dbuf<uint16_t> tensors[T_COUNT];
    int tensor_cp[T_COUNT];
    dbuf<float> out, target, grad_y[10], dx, wpartial, bpartial, block_loss, losses, radiance;
    dbuf<uint16_t> dy;
    dbuf<float> data_feat, data_target;
    dbuf<int32_t> ids;
    int32_t data_crops = 0, data_h = 0, data_w = 0;

    explicit ptd_network(uint64_t seed) {
        offsets[0] = 0;
        for(int l = 0; l < NLAYERS; l++) {
            offsets[l + 1] = offsets[l] + layer_params(SPECS[l]);
            cpa[l] = pad32(SPECS[l].cin_a);
            cpb[l] = SPECS[l].cin_b ? pad32(SPECS[l].cin_b) : 0;
            cpo[l] = pad32(SPECS[l].cout);
        }
        tensor_cp[T_X] = 32;
        for(int k = 0; k < 10; k++) tensor_cp[t_y(k)] = pad32(SPECS[k].cout);
        tensor_cp[T_P1] = tensor_cp[t_y(1)]; tensor_cp[T_P2] = tensor_cp[t_y(2)];
        tensor_cp[T_P3] = tensor_cp[t_y(3)]; tensor_cp[T_P4] = tensor_cp[t_y(4)];
        HC(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        {
            int device = 0;
            HC(hipGetDevice(&device));
            HC(hipDeviceGetAttribute(&compute_units, hipDeviceAttributeMultiprocessorCount, device));
        }
        HC(hipEventCreate(&ev0));
        HC(hipEventCreate(&ev1));
        const int64_t count = offsets[NLAYERS];
        params.ensure(size_t(count)); grads.ensure(size_t(count));
        adam_m.ensure(size_t(count)); adam_v.ensure(size_t(count));
        std::vector<float> init(size_t(count), 0.0f);
        // He-normal weights from a splitmix64 stream (Box-Muller); zero biases;
        // the output layer starts near zero so the network starts as the identity.
        uint64_t state = seed;
        auto next = [&]() {
            uint64_t z = (state += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto uniform = [&]() { return (double(next() >> 11) + 0.5) * 0x1p-53; };
        for(int l = 0; l < NLAYERS; l++) {
            const int cin = SPECS[l].cin_a + SPECS[l].cin_b;
            const double sd = l == NLAYERS - 1 ? 0.1 * std::sqrt(1.0 / (9.0 * cin)) : std::sqrt(2.0 / (9.0 * cin));
            const int64_t nw = int64_t(SPECS[l].cout) * 9 * cin;
            for(int64_t i = 0; i < nw; i++) {
                const double u1 = uniform(), u2 = uniform();
                init[size_t(offsets[l] + i)] = float(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
        }
        HC(hipMemcpy(params.p, init.data(), init.size() * sizeof(float), hipMemcpyHostToDevice));
        reset_optimizer();
        for(int l = 0; l < NLAYERS; l++) {
            wk[l].ensure(size_t(9) * size_t(cpa[l] + cpb[l]) * size_t(cpo[l]));
            wt[l].ensure(size_t(9) * size_t(cpa[l] + cpb[l]) * size_t(cpo[l]));
        }
        pack();
        HC(hipStreamSynchronize(stream));
    }
    ~ptd_network() {
        if(stream) (void)hipStreamSynchronize(stream);
        if(ev0) (void)hipEventDestroy(ev0);
        if(ev1) (void)hipEventDestroy(ev1);
        if(stream) (void)hipStreamDestroy(stream);
    }

    void reset_optimizer() {
        HC(hipMemsetAsync(adam_m.p, 0, adam_m.n * sizeof(float), stream));
        HC(hipMemsetAsync(adam_v.p, 0, adam_v.n * sizeof(float), stream));
        step = 0;
    }

    void pack() {
        for(int l = 0; l < NLAYERS; l++) {
            const layer_spec& s = SPECS[l];
            const int64_t count = int64_t(9) * (cpa[l] + cpb[l]) * cpo[l];
            hipLaunchKernelGGL(pack_forward_kernel, dim3(grid_for(count)), dim3(256), 0, stream, params.p + offsets[l],
                               s.cin_a, cpa[l], s.cin_b, cpb[l], s.cout, cpo[l], wk[l].p);
            hipLaunchKernelGGL(pack_backward_kernel, dim3(grid_for(count)), dim3(256), 0, stream, params.p + offsets[l],
                               s.cin_a, cpa[l], s.cin_b, cpb[l], s.cout, cpo[l], wt[l].p);
        }
        HC(hipGetLastError());
    }

    int level_h(int level) const { return h >> level; }
    int level_w(int level) const { return w >> level; }
    int64_t level_pixels(int level) const { return int64_t(n) * level_h(level) * level_w(level); }

    void shape(int nn, int hh, int ww, bool training) {
        if(nn <= 0 || hh <= 0 || ww <= 0 || hh % 16 || ww % 16) throw std::invalid_argument("denoiser shape must be positive multiples of 16");
        n = nn; h = hh; w = ww;
        auto level_of = [&](int t) {
            if(t == T_X) return 0;
            if(t >= T_P1) return t - T_P1 + 1;
            return SPECS[t - T_Y0].level;
        };
        for(int t = 0; t < T_COUNT; t++) tensors[t].ensure(size_t(level_pixels(level_of(t))) * size_t(tensor_cp[t]));
        out.ensure(size_t(level_pixels(0)) * 32);
        if(!training) return;
        target.ensure(size_t(level_pixels(0)) * 3);
        for(int k = 0; k < 10; k++) grad_y[k].ensure(size_t(level_pixels(SPECS[k].level)) * size_t(tensor_cp[t_y(k)]));
        size_t dx_max = 0, dy_max = 0;
        for(int l = 0; l < NLAYERS; l++) {
            dx_max = std::max(dx_max, size_t(level_pixels(SPECS[l].level)) * size_t(cpa[l] + cpb[l]));
            dy_max = std::max(dy_max, size_t(level_pixels(SPECS[l].level)) * size_t(cpo[l]));
        }
        dx.ensure(dx_max);
        dy.ensure(dy_max);
        size_t wp_max = 0, bp_max = 0;
        for(int l = 0; l < NLAYERS; l++) {
            int64_t splits, per_split;
            wgrad_plan(l, splits, per_split);
            wp_max = std::max(wp_max, size_t(splits) * 9 * size_t(cpa[l] + cpb[l]) * size_t(cpo[l]));
            bp_max = std::max(bp_max, size_t(splits) * size_t(cpo[l]));
        }
        wpartial.ensure(wp_max);
        bpartial.ensure(bp_max);
        block_loss.ensure(1024);
    }

    // Pixel splits of the weight-gradient reduction of layer l: (splits, sub-blocks per split).
    void wgrad_plan(int l, int64_t& splits, int64_t& per_split) const {
        const int cpi = cpa[l] + cpb[l];
        const int64_t nsub = (level_pixels(SPECS[l].level) + WG_PIX - 1) / WG_PIX;
        const int64_t want = std::max<int64_t>(1, 768 / (cpi / 32));
        per_split = (nsub + std::min(nsub, want) - 1) / std::min(nsub, want);
        splits = (nsub + per_split - 1) / per_split;
    }

    conv_src source(int l) const {
        const layer_spec& s = SPECS[l];
        conv_src c;
        c.a = tensors[s.src_a].p;
        c.b = s.src_b >= 0 ? tensors[s.src_b].p : nullptr;
        c.cpa = cpa[l]; c.cpb = cpb[l]; c.a_up = s.up_a ? 1 : 0;
        c.n = n; c.h = level_h(s.level); c.w = level_w(s.level);
        return c;
    }

    // The convolution's persistent grid: as many workgroups as stay resident
    // (conv_waves per compute unit, shared by the output-channel groups of
    // gridDim.y), fewer when there are fewer tiles, with the tiles dealt
    // evenly.  An EPI_F32 output of at most 16 channels computes one
    // 16-channel tile.
    template<int EPI>
    void launch_conv(const conv_src& src, const uint16_t *weights, int kchunks, int co_total, const float *bias, int cout,
                     void *dst, int frame_h = 0, int frame_w = 0, uint16_t *pooled = nullptr) {
        const uint16_t *xin = tensors[T_X].p;
        const int tiles = co_total / 16;
        const int tiles_x = (src.w + TILE - 1) / TILE, tiles_y = (src.h + TILE - 1) / TILE;
        const int64_t blocks = int64_t(src.n) * tiles_x * tiles_y;
        if(blocks > INT32_MAX || int64_t(src.h) * src.w * std::max(src.cpa, src.cpb) >= INT32_MAX)
            throw std::invalid_argument("denoiser input too large");
        // workgroups per image-tile column (gy columns): all resident at once, tiles dealt evenly
        auto grid_x = [&](int mt, int gy) {
            const int64_t resident = int64_t(compute_units) * (mt <= 2 ? conv_waves<2> : conv_waves<4>) /
                                     std::max(1, gy);
            const int64_t per_block = (blocks + std::max<int64_t>(1, resident) - 1) / std::max<int64_t>(1, resident);
            return unsigned((blocks + per_block - 1) / per_block);
        };
        if((EPI == EPI_F32 || EPI == EPI_RADIANCE) && cout <= 16)
            hipLaunchKernelGGL((conv_tile_kernel<1, EPI>), dim3(grid_x(1, 1), 1u), dim3(256), 0, stream, src, weights,
                               kchunks, co_total, bias, cout, dst, tiles_x, tiles_y, xin, frame_h, frame_w, pooled);
        else if(tiles % 6 == 0)
            hipLaunchKernelGGL((conv_tile_kernel<6, EPI>), dim3(grid_x(6, tiles / 6), unsigned(tiles / 6)), dim3(256), 0, stream,
                               src, weights, kchunks, co_total, bias, cout, dst, tiles_x, tiles_y, xin, frame_h, frame_w, pooled);
        else if(tiles % 4 == 0)
            hipLaunchKernelGGL((conv_tile_kernel<4, EPI>), dim3(grid_x(4, tiles / 4), unsigned(tiles / 4)), dim3(256), 0, stream,
                               src, weights, kchunks, co_total, bias, cout, dst, tiles_x, tiles_y, xin, frame_h, frame_w, pooled);
        else
            hipLaunchKernelGGL((conv_tile_kernel<2, EPI>), dim3(grid_x(2, tiles / 2), unsigned(tiles / 2)), dim3(256), 0, stream,
                               src, weights, kchunks, co_total, bias, cout, dst, tiles_x, tiles_y, xin, frame_h, frame_w, pooled);
        HC(hipGetLastError());
    }

    // The forward pass; with `radiance`, the output layer writes the denoised
    // radiance of the frame_h x frame_w frame (inference) instead of `out`.
    void forward(float *radiance = nullptr, int frame_h = 0, int frame_w = 0) {
        for(int l = 0; l < NLAYERS; l++) {
            const layer_spec& s = SPECS[l];
            const conv_src src = source(l);
            const float *bias = params.p + offsets[l] + int64_t(s.cout) * 9 * (s.cin_a + s.cin_b);
            const int kchunks = (cpa[l] + cpb[l]) / 32;
            if(l == NLAYERS - 1 && radiance)
                launch_conv<EPI_RADIANCE>(src, wk[l].p, kchunks, cpo[l], bias, s.cout, radiance, frame_h, frame_w);
            else if(l == NLAYERS - 1) launch_conv<EPI_F32>(src, wk[l].p, kchunks, cpo[l], bias, s.cout, out.p);
            else  // encoder layers 1..4 also write the 2x2 average that the next level reads
                launch_conv<EPI_BF16_RELU>(src, wk[l].p, kchunks, cpo[l], bias, s.cout, tensors[t_y(l)].p, 0, 0,
                                           l >= 1 && l <= 4 ? tensors[T_P1 + (l - 1)].p : nullptr);
        }
    }

    // Loss gradient in dy (for the output layer); per-step mean loss into losses[slot].
    void loss(int slot) {
        const int64_t pixels = level_pixels(0);
        const unsigned blocks = unsigned(std::min<int64_t>(1024, (pixels + 255) / 256));
        hipLaunchKernelGGL(loss_kernel, dim3(blocks), dim3(256), 0, stream, out.p, tensors[T_X].p, target.p, pixels, dy.p,
                           block_loss.p);
        hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(64), 0, stream, block_loss.p, int(blocks), losses.p + slot);
        HC(hipGetLastError());
    }

    void weight_gradient(int l) {
        const layer_spec& s = SPECS[l];
        const conv_src src = source(l);
        const int cpi = cpa[l] + cpb[l];
        int64_t used, bps;
        wgrad_plan(l, used, bps);
        const dim3 grid(unsigned(used), unsigned(cpi / 32));
        switch(cpo[l] / 16) {
            case 2: hipLaunchKernelGGL(wgrad_kernel<2>, grid, dim3(256), 0, stream, src, dy.p, cpo[l], bps, cpi, wpartial.p, bpartial.p); break;
            case 4: hipLaunchKernelGGL(wgrad_kernel<4>, grid, dim3(256), 0, stream, src, dy.p, cpo[l], bps, cpi, wpartial.p, bpartial.p); break;
            case 6: hipLaunchKernelGGL(wgrad_kernel<6>, grid, dim3(256), 0, stream, src, dy.p, cpo[l], bps, cpi, wpartial.p, bpartial.p); break;
            default: throw std::logic_error("unsupported output width");
        }
        const int64_t nw = int64_t(s.cout) * 9 * (s.cin_a + s.cin_b);
        const int64_t per_split = int64_t(9) * cpi * cpo[l];
        hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(unsigned((per_split + 63) / 64)), dim3(64, 4), 0, stream, wpartial.p,
                           int(used), cpi, cpo[l], s.cin_a, cpa[l], s.cin_b, s.cout, grads.p + offsets[l]);
        hipLaunchKernelGGL(bgrad_reduce_kernel, dim3(unsigned(s.cout)), dim3(256), 0, stream, bpartial.p, int(used),
                           cpo[l], s.cout, grads.p + offsets[l] + nw);
        HC(hipGetLastError());
    }

    // Backward through all layers from dy = dL/d(out); fills grads.
    void backward() {
        // grad_y[k] is written by its first contribution (consumers come
        // later in the layer order, so backward reaches them first) and added
        // to by the others; every hidden output has at least one consumer.
        bool written[10] = {};
        auto first = [&](int k) { const bool f = !written[k]; written[k] = true; return f ? 1 : 0; };
        for(int l = NLAYERS - 1; l >= 0; l--) {
            if(l < NLAYERS - 1 && !written[l]) throw std::logic_error("hidden output without a gradient");
            const layer_spec& s = SPECS[l];
            if(l < NLAYERS - 1) {  // dy of this layer from its output's gradient and ReLU
                const int64_t count = level_pixels(s.level) * tensor_cp[t_y(l)];
                hipLaunchKernelGGL(relu_grad_kernel, dim3(grid_for(count / 8)), dim3(256), 0, stream, dy.p, grad_y[l].p,
                                   tensors[t_y(l)].p, count);
            }
            weight_gradient(l);
            if(l == 0) break;
            // dX of this layer's (virtual) input.
            const int cpi = cpa[l] + cpb[l];
            conv_src g;
            g.a = dy.p; g.b = nullptr; g.cpa = cpo[l]; g.cpb = 0; g.a_up = 0;
            g.n = n; g.h = level_h(s.level); g.w = level_w(s.level);
            launch_conv<EPI_F32>(g, wt[l].p, cpo[l] / 32, cpi, nullptr, cpi, dx.p);  // every channel of dX stored
            const int64_t pixels = level_pixels(s.level);
            if(s.src_b >= 0) {  // concat: upsampled deeper output, then the skip tensor
                const int ka = s.src_a - T_Y0, kb = s.src_b - T_Y0;
                const int64_t ca = level_pixels(s.level + 1) * cpa[l] / 4;
                hipLaunchKernelGGL(route_up_kernel, dim3(grid_for(ca)), dim3(256), 0, stream, grad_y[ka].p, dx.p, n,
                                   level_h(s.level), level_w(s.level), cpa[l], cpi, first(ka));
                const int64_t cb = pixels * cpb[l] / 4;
                hipLaunchKernelGGL(route_same_kernel, dim3(grid_for(cb)), dim3(256), 0, stream, grad_y[kb].p, dx.p, pixels,
                                   cpb[l], cpi, cpa[l], first(kb));
            } else if(s.src_a >= T_P1) {  // pooled encoder output
                const int k = l - 1;
                const int lv = SPECS[k].level;
                const int64_t count = level_pixels(lv + 1) * tensor_cp[t_y(k)] / 4;
                hipLaunchKernelGGL(route_pool_kernel, dim3(grid_for(count)), dim3(256), 0, stream, grad_y[k].p, dx.p, n,
                                   level_h(lv), level_w(lv), tensor_cp[t_y(k)], first(k));
            } else {  // plain: the previous layer's output
                const int k = s.src_a - T_Y0;
                const int64_t count = pixels * cpi / 4;
                hipLaunchKernelGGL(route_same_kernel, dim3(grid_for(count)), dim3(256), 0, stream, grad_y[k].p, dx.p, pixels,
                                   cpi, cpi, 0, first(k));
            }
            HC(hipGetLastError());
        }
    }

    void adam(float lr) {
        step++;
        const float b1 = 0.9f, b2 = 0.99f;
        const float c1 = float(1.0 - std::pow(double(b1), double(step))), c2 = float(1.0 - std::pow(double(b2), double(step)));
        const int64_t count = offsets[NLAYERS];
        hipLaunchKernelGGL(adam_kernel, dim3(grid_for(count)), dim3(256), 0, stream, params.p, grads.p, adam_m.p, adam_v.p,
                           count, lr, b1, b2, 1e-8f, c1, c2);
        HC(hipGetLastError());
        pack();
    }
};

namespace {

thread_local std::string ptd_error;

template<class F> pt_status ptd_guard(F&& f) {
    try {
        f();
        return PT_OK;
    } catch(const std::invalid_argument& e) {
        ptd_error = e.what();
        return PT_EINVAL;
    } catch(const std::exception& e) {
        ptd_error = e.what();
        return PT_EDEVICE;
    }
}

} // namespace

extern "C" {

const char *ptd_last_error(void) { return ptd_error.c_str(); }

int64_t ptd_parameter_count(void) {
    int64_t count = 0;
    for(int l = 0; l < NLAYERS; l++) count += layer_params(SPECS[l]);
    return count;
}

pt_status ptd_create(uint64_t seed, ptd_network **out) {
    return ptd_guard([&] {
        if(!out) throw std::invalid_argument("null argument");
        *out = new ptd_network(seed);
    });
}

void ptd_destroy(ptd_network *net) { delete net; }

pt_status ptd_get_parameters(ptd_network *net, float *host, int64_t count) {
    return ptd_guard([&] {
        if(!net || !host || count != net->offsets[NLAYERS]) throw std::invalid_argument("parameter buffer size mismatch");
        HC(hipStreamSynchronize(net->stream));
        HC(hipMemcpy(host, net->params.p, size_t(count) * sizeof(float), hipMemcpyDeviceToHost));
    });
}

pt_status ptd_set_parameters(ptd_network *net, const float *host, int64_t count) {
    return ptd_guard([&] {
        if(!net || !host || count != net->offsets[NLAYERS]) throw std::invalid_argument("parameter buffer size mismatch");
        HC(hipStreamSynchronize(net->stream));
        HC(hipMemcpy(net->params.p, host, size_t(count) * sizeof(float), hipMemcpyHostToDevice));
        net->pack();
        net->reset_optimizer();
        HC(hipStreamSynchronize(net->stream));
    });
}

pt_status ptd_forward(ptd_network *net, const float *features, int32_t n, int32_t h, int32_t w, float *radiance) {
    return ptd_guard([&] {
        if(!net || !features || !radiance) throw std::invalid_argument("null argument");
        net->shape(n, h, w, false);
        const int64_t pixels = int64_t(n) * h * w;
        net->radiance.ensure(size_t(pixels) * 3);
        dbuf<float> feat;
        feat.ensure(size_t(pixels) * 9);
        HC(hipMemcpyAsync(feat.p, features, size_t(pixels) * 9 * sizeof(float), hipMemcpyHostToDevice, net->stream));
        hipLaunchKernelGGL(features_kernel, dim3(grid_for(pixels)), dim3(256), 0, net->stream, feat.p, pixels,
                           net->tensors[T_X].p);
        net->forward(net->radiance.p, h, w);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(radiance, net->radiance.p, size_t(pixels) * 3 * sizeof(float), hipMemcpyDeviceToHost, net->stream));
        HC(hipStreamSynchronize(net->stream));
    });
}

pt_status ptd_loss_and_gradients(ptd_network *net, const float *features, const float *targets, int32_t n, int32_t h,
                                 int32_t w, float *loss, float *gradients, int64_t count) {
    return ptd_guard([&] {
        if(!net || !features || !targets || !loss) throw std::invalid_argument("null argument");
        if(gradients && count != net->offsets[NLAYERS]) throw std::invalid_argument("gradient buffer size mismatch");
        net->shape(n, h, w, true);
        const int64_t pixels = int64_t(n) * h * w;
        net->losses.ensure(1);
        dbuf<float> feat;
        feat.ensure(size_t(pixels) * 9);
        HC(hipMemcpyAsync(feat.p, features, size_t(pixels) * 9 * sizeof(float), hipMemcpyHostToDevice, net->stream));
        HC(hipMemcpyAsync(net->target.p, targets, size_t(pixels) * 3 * sizeof(float), hipMemcpyHostToDevice, net->stream));
        hipLaunchKernelGGL(features_kernel, dim3(grid_for(pixels)), dim3(256), 0, net->stream, feat.p, pixels,
                           net->tensors[T_X].p);
        net->forward();
        net->loss(0);
        net->backward();
        HC(hipMemcpyAsync(loss, net->losses.p, sizeof(float), hipMemcpyDeviceToHost, net->stream));
        if(gradients)
            HC(hipMemcpyAsync(gradients, net->grads.p, size_t(count) * sizeof(float), hipMemcpyDeviceToHost, net->stream));
        HC(hipStreamSynchronize(net->stream));
    });
}

pt_status ptd_set_training_data(ptd_network *net, const float *features, const float *targets, int32_t crops, int32_t h,
                                int32_t w) {
    return ptd_guard([&] {
        if(!net || !features || !targets || crops <= 0 || h <= 0 || w <= 0) throw std::invalid_argument("invalid training data");
        const size_t pixels = size_t(crops) * size_t(h) * size_t(w);
        net->data_feat.ensure(pixels * 9);
        net->data_target.ensure(pixels * 3);
        HC(hipMemcpy(net->data_feat.p, features, pixels * 9 * sizeof(float), hipMemcpyHostToDevice));
        HC(hipMemcpy(net->data_target.p, targets, pixels * 3 * sizeof(float), hipMemcpyHostToDevice));
        net->data_crops = crops; net->data_h = h; net->data_w = w;
    });
}

pt_status ptd_train(ptd_network *net, const int32_t *crop_ids, int32_t batch, int32_t steps, float learning_rate,
                    float *losses, float *milliseconds) {
    return ptd_guard([&] {
        if(!net || !crop_ids || batch <= 0 || steps <= 0) throw std::invalid_argument("invalid training arguments");
        if(!net->data_crops) throw std::invalid_argument("no training data");
        for(int64_t i = 0; i < int64_t(batch) * steps; i++)
            if(crop_ids[i] < 0 || crop_ids[i] >= net->data_crops) throw std::invalid_argument("crop id out of range");
        net->shape(batch, net->data_h, net->data_w, true);
        net->ids.ensure(size_t(batch) * size_t(steps));
        net->losses.ensure(size_t(steps));
        HC(hipMemcpyAsync(net->ids.p, crop_ids, size_t(batch) * size_t(steps) * sizeof(int32_t), hipMemcpyHostToDevice,
                          net->stream));
        const int64_t hw = int64_t(net->data_h) * net->data_w;
        HC(hipEventRecord(net->ev0, net->stream));
        for(int s = 0; s < steps; s++) {
            hipLaunchKernelGGL(gather_kernel, dim3(grid_for(hw * batch)), dim3(256), 0, net->stream, net->data_feat.p,
                               net->data_target.p, net->ids.p + size_t(s) * size_t(batch), batch, hw, net->tensors[T_X].p,
                               net->target.p);
            net->forward();
            net->loss(s);
            net->backward();
            net->adam(learning_rate);
        }
        HC(hipEventRecord(net->ev1, net->stream));
        if(losses) HC(hipMemcpyAsync(losses, net->losses.p, size_t(steps) * sizeof(float), hipMemcpyDeviceToHost, net->stream));
        HC(hipStreamSynchronize(net->stream));
        float ms = 0;
        HC(hipEventElapsedTime(&ms, net->ev0, net->ev1));
        if(milliseconds) *milliseconds = ms;
    });
}

pt_status ptd_denoise_renderer(ptd_network *net, const pt_renderer *renderer, float *radiance, float *milliseconds) {
    return ptd_guard([&] {
        if(!net || !renderer || !radiance) throw std::invalid_argument("null argument");
        pt::frame_view f;
        if(!pt::renderer_frame(renderer, &f)) throw std::invalid_argument("invalid renderer");
        if(!f.aov) throw std::invalid_argument("the renderer's feature buffers are not enabled");
        if(f.samples <= 0) throw std::invalid_argument("no samples accumulated");
        HC(hipStreamSynchronize(f.stream));
        const int hp = (f.height + 15) / 16 * 16, wp = (f.width + 15) / 16 * 16;
        net->shape(1, hp, wp, false);
        const int64_t pixels = int64_t(f.width) * f.height;
        net->radiance.ensure(size_t(pixels) * 3);
        HC(hipStreamSynchronize(net->stream));
        HC(hipEventRecord(net->ev0, net->stream));
        hipLaunchKernelGGL(frame_features_kernel, dim3(grid_for(int64_t(hp) * wp)), dim3(256), 0, net->stream, f.sum, f.aov,
                           float(1.0 / double(f.samples)), f.width, f.height, wp, hp, net->tensors[T_X].p);
        net->forward(net->radiance.p, f.height, f.width);
        HC(hipGetLastError());
        HC(hipEventRecord(net->ev1, net->stream));
        HC(hipMemcpyAsync(radiance, net->radiance.p, size_t(pixels) * 3 * sizeof(float), hipMemcpyDeviceToHost, net->stream));
        HC(hipStreamSynchronize(net->stream));
        float ms = 0;
        HC(hipEventElapsedTime(&ms, net->ev0, net->ev1));
        if(milliseconds) *milliseconds = ms;
    });
}

}
