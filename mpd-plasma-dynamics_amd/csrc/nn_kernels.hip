// nn_kernels.hip -- gfx950 kernels of the learned warm start: a BF16 conv
// encoder-decoder (3x3 convolutions, FP32 accumulate on the matrix cores)
// with its training kernels.  C ABI in include/mpd_nn.h; every pointer is a
// device pointer, every tensor NHWC (batch, z rows, r columns, channels),
// activations and MFMA operands bfloat16, pre-activations and gradients FP32.
//
// The three convolution passes are one implicit GEMM C[m][n] = sum_k A(m,k)
// B(n,k) with mode-specific operand gathers:
//   forward      m = pixel, n = output channel, k = (tap, input channel)
//   data grad    m = pixel, n = input channel,  k = (tap, output channel)
//   weight grad  m = output channel, n = (tap, input channel), k = pixel
//                (split over pixel ranges; partial sums reduced in a fixed
//                order, so the result is deterministic)
// conv_gemm (any channel count): a 256-thread workgroup computes a 64 x 64
// tile of C from element-staged 64 x 32 operand tiles, each wave a 32 x 32
// sub-tile as 2 x 2 v_mfma_f32_16x16x16_bf16 accumulators.  conv_gemm2
// (channel counts multiples of 8): tiles up to 128 x 128 on
// v_mfma_f32_16x16x32_bf16 from 16-byte runs.  wgrad_lds (the weight
// gradient, channel counts multiples of 16): row blocks staged in LDS and
// read with gfx950's LDS transpose reads.  Fused passes for the training
// step: the forward GEMM writes SiLU's output beside z, the SiLU backward
// writes dz with the bias gradient's slices (for a layer whose output feeds
// the next conv directly, in that conv's data-gradient epilogue), the
// decoder's upsample +
// concatenation (and its backward) and the encoder's pooling backward +
// skip gradient are one pass each, Adam runs over every tensor at once;
// each equals its separate passes bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/mpd_nn.h"

namespace {

thread_local std::string g_nn_error;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even (finite inputs; NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }

// SiLU's output as bf16, and its backward dz = da silu'(z), silu'(z) =
// s (1 + z (1 - s)) with s = sigmoid(z): one definition for the separate
// element-wise kernels and the ones fused into other passes.
__device__ __forceinline__ uint16_t silu_bf16(float v) { return f2bf(v * sigmoid(v)); }
__device__ __forceinline__ float silu_grad(float v, float da) {
    const float s = sigmoid(v);
    return da * (s * (1.f + v * (1.f - s)));
}

enum { MODE_FWD = 0, MODE_DGRAD = 1, MODE_WGRAD = 2 };

struct ConvArgs {
    const uint16_t* x;   // input activations [B][H][W][Cin]
    const uint16_t* w;   // weights [Cout][3][3][Cin]
    const uint16_t* dz;  // pre-activation gradient [B][H][W][Cout]
    const float* bias;   // [Cout] (forward)
    float* out;          // forward: z [P][Cout]; data grad: dx [P][Cin]; weight grad: partials [S][Cout][9 Cin]
    uint16_t* act;       // forward: bf16(silu(z)) [P][Cout] as well, or null
    // Data gradient with the SiLU backward of the layer below fused (zact
    // non-null): dzo = bf16(out * silu'(zact)) [P][Cin] instead of out, and
    // per pixel tile column sums of those values into bpart [tiles][Cin].
    const float* zact;
    uint16_t* dzo;
    float* bpart;
    int B, H, W, Cin, Cout;
    int M, N, K;
    int kchunk;  // weight grad: pixels per split
};

constexpr int BM = 64, BN = 64, BK = 32, LDP = BK + 8;  // LDS row pitch (bf16)

// Input activation at pixel m shifted by tap (ky, kx) - 1, channel ci; zero
// outside the image (same padding).
__device__ __forceinline__ uint16_t gather_x(const ConvArgs& a, int m, int tap, int ci) {
    const int hw = a.H * a.W;
    const int b = m / hw, r = m - b * hw;
    const int y = r / a.W + tap / 3 - 1, x = r % a.W + tap % 3 - 1;
    if (y < 0 || y >= a.H || x < 0 || x >= a.W) return 0;
    return a.x[((int64_t)(b * a.H + y) * a.W + x) * a.Cin + ci];
}

template <int MODE>
__device__ __forceinline__ uint16_t load_a(const ConvArgs& a, int m, int k) {
    if (m >= a.M || k >= a.K) return 0;
    if (MODE == MODE_FWD) {
        const int tap = k / a.Cin, ci = k - tap * a.Cin;
        return gather_x(a, m, tap, ci);
    } else if (MODE == MODE_DGRAD) {
        // dx[y][x][ci] = sum dz[y + 1 - ky][x + 1 - kx][co] w[co][ky][kx][ci]
        const int tap = k / a.Cout, co = k - tap * a.Cout;
        const int hw = a.H * a.W;
        const int b = m / hw, r = m - b * hw;
        const int y = r / a.W + 1 - tap / 3, x = r % a.W + 1 - tap % 3;
        if (y < 0 || y >= a.H || x < 0 || x >= a.W) return 0;
        return a.dz[((int64_t)(b * a.H + y) * a.W + x) * a.Cout + co];
    } else {
        // A(co, pixel) = dz[pixel][co]
        return a.dz[(int64_t)k * a.Cout + m];
    }
}

template <int MODE>
__device__ __forceinline__ uint16_t load_b(const ConvArgs& a, int n, int k) {
    if (n >= a.N || k >= a.K) return 0;
    if (MODE == MODE_FWD) {
        return a.w[(int64_t)n * a.K + k];
    } else if (MODE == MODE_DGRAD) {
        const int tap = k / a.Cout, co = k - tap * a.Cout;
        return a.w[((int64_t)co * 9 + tap) * a.Cin + n];
    } else {
        // B((tap, ci), pixel) = input at pixel shifted by tap
        const int tap = n / a.Cin, ci = n - tap * a.Cin;
        return gather_x(a, k, tap, ci);
    }
}

// Staging order of the operand tiles: element e of a 64 x 32 tile is (row
// e / 32, column e % 32) when the source is contiguous along k (consecutive
// threads read consecutive k), (row e % 64, column e / 64) when it is
// contiguous along the tile's rows (consecutive output channels / input
// channels).  Forward: A and B contiguous in k.  Data gradient: A (dz) in k,
// B (weights, n = input channel) in rows.  Weight gradient: A (dz, m =
// output channel) and B (input, n = (tap, input channel)) in rows.
template <int MODE>
__device__ __forceinline__ constexpr bool a_rows_fast() { return MODE == MODE_WGRAD; }
template <int MODE>
__device__ __forceinline__ constexpr bool b_rows_fast() { return MODE != MODE_FWD; }

// Vector staging (every channel count a multiple of 8): each thread moves one
// 8-element (16-byte) run of each operand tile per chunk.  The run lies along
// k when the source is contiguous in k, else along the tile's rows; runs
// never cross a tap (channel counts and run starts are multiples of 8), so a
// run is wholly inside or wholly outside the image and the tile.
union Run8 {
    uint4 v;
    uint16_t h[8];
};

__device__ __forceinline__ uint4 zero_run() { return uint4{0u, 0u, 0u, 0u}; }

// The pixel of row m shifted by (dy, dx), or -1 outside the image.
__device__ __forceinline__ int64_t shifted_pixel(const ConvArgs& a, int m, int dy, int dx) {
    const int hw = a.H * a.W;
    const int b = m / hw, r = m - b * hw;
    const int y = r / a.W + dy, x = r % a.W + dx;
    if (y < 0 || y >= a.H || x < 0 || x >= a.W) return -1;
    return (int64_t)(b * a.H + y) * a.W + x;
}

template <int MODE>
__device__ __forceinline__ uint4 vload_a(const ConvArgs& a, int m, int k) {
    if (m >= a.M || k >= a.K) return zero_run();
    if (MODE == MODE_FWD) {
        const int tap = k / a.Cin, ci = k - tap * a.Cin;
        const int64_t p = shifted_pixel(a, m, tap / 3 - 1, tap % 3 - 1);
        return p < 0 ? zero_run() : *reinterpret_cast<const uint4*>(a.x + p * a.Cin + ci);
    } else if (MODE == MODE_DGRAD) {
        const int tap = k / a.Cout, co = k - tap * a.Cout;
        const int64_t p = shifted_pixel(a, m, 1 - tap / 3, 1 - tap % 3);
        return p < 0 ? zero_run() : *reinterpret_cast<const uint4*>(a.dz + p * a.Cout + co);
    } else {
        // rows m .. m + 7 (output channels) of pixel k
        return *reinterpret_cast<const uint4*>(a.dz + (int64_t)k * a.Cout + m);
    }
}

template <int MODE>
__device__ __forceinline__ uint4 vload_b(const ConvArgs& a, int n, int k) {
    if (n >= a.N || k >= a.K) return zero_run();
    if (MODE == MODE_FWD) {
        return *reinterpret_cast<const uint4*>(a.w + (int64_t)n * a.K + k);
    } else if (MODE == MODE_DGRAD) {
        // rows n .. n + 7 (input channels) of weight column (tap, co)
        const int tap = k / a.Cout, co = k - tap * a.Cout;
        return *reinterpret_cast<const uint4*>(a.w + ((int64_t)co * 9 + tap) * a.Cin + n);
    } else {
        // rows n .. n + 7 = (tap, ci .. ci + 7) of pixel k
        const int tap = n / a.Cin, ci = n - tap * a.Cin;
        const int64_t p = shifted_pixel(a, k, tap / 3 - 1, tap % 3 - 1);
        return p < 0 ? zero_run() : *reinterpret_cast<const uint4*>(a.x + p * a.Cin + ci);
    }
}

// Thread t's run of a 64 x 32 tile: along k, row t / 4 and columns
// 8 (t % 4) .. + 7; along rows, rows 8 (t % 8) .. + 7 and column t / 8.
template <bool kRowsFast, class Load>
__device__ __forceinline__ uint4 stage_load(int t, int r0, int k0, int k_end, Load ld) {
    if (kRowsFast) {
        const int k = k0 + (t >> 3);
        return (k < k_end) ? ld(r0 + 8 * (t & 7), k) : zero_run();
    }
    const int k = k0 + 8 * (t & 3);
    return (k < k_end) ? ld(r0 + (t >> 2), k) : zero_run();
}

template <bool kRowsFast>
__device__ __forceinline__ void stage_store(int t, uint16_t* S, uint4 v) {
    if (kRowsFast) {
        Run8 u;
        u.v = v;
        const int c = t >> 3, r = 8 * (t & 7);
#pragma unroll
        for (int j = 0; j < 8; ++j) S[(r + j) * LDP + c] = u.h[j];
    } else {
        *reinterpret_cast<uint4*>(&S[(t >> 2) * LDP + 8 * (t & 3)]) = v;
    }
}

template <int MODE, bool VEC>
__global__ void __launch_bounds__(256) conv_gemm(ConvArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDP];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LDP];
    const int t = (int)threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = (int)blockIdx.x * BM, n0 = (int)blockIdx.y * BN;
    int k_begin = 0, k_end = a.K;
    if (MODE == MODE_WGRAD) {
        k_begin = (int)blockIdx.z * a.kchunk;
        k_end = min(a.K, k_begin + a.kchunk);
    }
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int row = lane & 15, kq = (lane >> 4) * 4;
    for (int k0 = k_begin; k0 < k_end; k0 += BK) {
        if (VEC) {
            const uint4 ra = stage_load<a_rows_fast<MODE>()>(t, m0, k0, k_end,
                                                              [&](int m, int k) { return vload_a<MODE>(a, m, k); });
            const uint4 rb = stage_load<b_rows_fast<MODE>()>(t, n0, k0, k_end,
                                                              [&](int n, int k) { return vload_b<MODE>(a, n, k); });
            stage_store<a_rows_fast<MODE>()>(t, As, ra);
            stage_store<b_rows_fast<MODE>()>(t, Bs, rb);
        } else {
        // Stage the 64 x 32 operand tiles element by element (e = t + 256 j).
        uint16_t va[8], vb[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = t + 256 * j;
            const int ra = a_rows_fast<MODE>() ? (e & 63) : (e >> 5), ca = a_rows_fast<MODE>() ? (e >> 6) : (e & 31);
            const int rb = b_rows_fast<MODE>() ? (e & 63) : (e >> 5), cb = b_rows_fast<MODE>() ? (e >> 6) : (e & 31);
            va[j] = (k0 + ca < k_end) ? load_a<MODE>(a, m0 + ra, k0 + ca) : (uint16_t)0;
            vb[j] = (k0 + cb < k_end) ? load_b<MODE>(a, n0 + rb, k0 + cb) : (uint16_t)0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int e = t + 256 * j;
            const int ra = a_rows_fast<MODE>() ? (e & 63) : (e >> 5), ca = a_rows_fast<MODE>() ? (e >> 6) : (e & 31);
            const int rb = b_rows_fast<MODE>() ? (e & 63) : (e >> 5), cb = b_rows_fast<MODE>() ? (e >> 6) : (e & 31);
            As[ra * LDP + ca] = va[j];
            Bs[rb * LDP + cb] = vb[j];
        }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf16x4 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint16_t* pa = &As[(wm * 32 + i * 16 + row) * LDP + ks + kq];
                const uint16_t* pb = &Bs[(wn * 32 + i * 16 + row) * LDP + ks + kq];
                fa[i] = *reinterpret_cast<const bf16x4*>(pa);
                fb[i] = *reinterpret_cast<const bf16x4*>(pb);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // Lane holds C[4 (lane / 16) + jj][lane % 16] of each 16 x 16 sub-tile.
    float* out = a.out;
    if (MODE == MODE_WGRAD) out += (int64_t)blockIdx.z * a.M * a.N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int m = m0 + wm * 32 + i * 16 + kq + jj;
                const int n = n0 + wn * 32 + j * 16 + row;
                if (m < a.M && n < a.N) {
                    float v = acc[i][j][jj];
                    if (MODE == MODE_FWD && a.bias) v += a.bias[n];
                    out[(int64_t)m * a.N + n] = v;
                    if (MODE == MODE_FWD && a.act) a.act[(int64_t)m * a.N + n] = silu_bf16(v);
                }
            }
}

// The same implicit GEMM for channel counts that are multiples of 8, on
// gfx950's v_mfma_f32_16x16x32_bf16 (K = 32 per instruction): a 256-thread
// workgroup computes a TM x TN tile of C (up to 128 x 128; the narrow
// shapes serve the output head and the small channel counts), the four
// waves in a WM x WN arrangement, each a (TM / WM) x (TN / WN) sub-tile of
// 16 x 16 accumulators.  Operand tiles of TM x 32 and TN x 32 go through
// two LDS buffers: the global loads of k-tile j + 1 are issued before the
// matrix instructions of tile j and stored after them, into the other
// buffer, so one barrier per k-tile separates the two.  Lane l of an MFMA
// holds row l % 16 and the 8 k values 8 (l / 16) .. + 7 of each operand;
// its results are C[4 (l / 16) + i][l % 16], i < 4.  An operand whose
// source is contiguous along k is staged row by row (a tile row holds its 32
// k values) and read back as one 16-byte run per lane; one contiguous along
// its rows is staged in source order (a tile row holds one k value for the
// T rows) and read with gfx950's LDS transpose read (below).  Both stage
// every 16-byte run with one 16-byte store.
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
constexpr int BK2 = 32, LDK2 = BK2 + 8;  // LDS row pitch (bf16) of a k-contiguous tile: 80 bytes

__device__ __forceinline__ bf16x4 tr_read(const uint16_t* p) {
    typedef __attribute__((address_space(3))) bf16x4 lds_v4;
    return __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4*)p);
}

__device__ __forceinline__ bf16x8v join8(bf16x4 lo, bf16x4 hi) {
    return __builtin_bit_cast(bf16x8v, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

// Row pitch (bf16) of a tile staged in source order (32 k rows of T
// values): half the pitch in 4-byte words is 8 or 24 mod 32, so the four k
// rows one 16-lane group of a transpose read touches (8 words each) fall on
// distinct banks.
template <int T>
__host__ __device__ constexpr int kt_pitch() { return T <= 32 ? 48 : T + 16; }
template <int T, bool kRowsFast>
__host__ __device__ constexpr int tile_elems() { return kRowsFast ? BK2 * kt_pitch<T>() : T * LDK2; }

// Run r of a T x 32 operand tile: along k, tile row r / 4, k columns
// 8 (r % 4) .. + 7; along rows, rows 8 (r % (T / 8)) .. + 7 of k column
// r / (T / 8), stored at that k's row of the source-order tile.
template <int T, bool kRowsFast>
__device__ __forceinline__ void stage_store2(int r, uint16_t* S, uint4 v) {
    if (kRowsFast)
        *reinterpret_cast<uint4*>(&S[(r / (T / 8)) * kt_pitch<T>() + 8 * (r % (T / 8))]) = v;
    else
        *reinterpret_cast<uint4*>(&S[(r >> 2) * LDK2 + 8 * (r & 3)]) = v;
}

// The MFMA operand of rows rb .. rb + 15 of a staged tile for this lane
// (row l % 16, k values 8 (l / 16) .. + 7).  Source-order tile: in each
// 16-lane group g, lane s reads k row 8 g + s / 4 (then + 4) at rows
// rb + 4 (s % 4) .. + 3, and the transpose read hands output lane 4 a + b
// element b of the reads of lanes a, a + 4, a + 8, a + 12: k rows
// 8 g .. 8 g + 3 (then + 4 .. + 7) of row rb + 4 a + b.
template <int T, bool kRowsFast>
__device__ __forceinline__ bf16x8v frag2(const uint16_t* S, int rb, int lane) {
    const int s = lane & 15, g = lane >> 4;
    if (kRowsFast) {
        constexpr int P = kt_pitch<T>();
        const uint16_t* p0 = S + (8 * g + (s >> 2)) * P + rb + 4 * (s & 3);
        return join8(tr_read(p0), tr_read(p0 + 4 * P));
    }
    return __builtin_bit_cast(bf16x8v, *reinterpret_cast<const uint4*>(&S[(rb + s) * LDK2 + 8 * g]));
}

// The data gradient's epilogue with the SiLU backward of the layer below,
// one row of waves at a time through an LDS slab (as the forward's): each
// thread takes 4-float runs of output rows (its column group fixed, 256 being
// a multiple of the TN / 4 groups), forms d = dx silu'(z) from whole 16-byte
// pieces of z, writes bf16 d, and adds d to its own 4 column sums; the sums
// of the threads of one column group are then added in thread order, and the
// tile's column sums go to bpart[tile_m] -- a deterministic order.
template <int WM, int TM, int TN>
__device__ __forceinline__ void dgrad_silu_rows(const ConvArgs& a, float* slab, int m0, int n0, int h, float4& cs) {
    constexpr int SP = TN + 4, SR = TM / WM, CG = TN / 4;
    constexpr int RUNS = SR * CG, RPT = (RUNS + 255) / 256;
    const int t = (int)threadIdx.x;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = t + 256 * q;
        if (RUNS % 256 == 0 || r < RUNS) {
            const int sr = r / CG, c4 = 4 * (r % CG);
            const int m = m0 + h * SR + sr, n = n0 + c4;
            if (m < a.M && n < a.N) {
                const int64_t e = (int64_t)m * a.N + n;
                const float4 v = *reinterpret_cast<const float4*>(&slab[sr * SP + c4]);
                const float4 z = *reinterpret_cast<const float4*>(a.zact + e);
                const float d0 = silu_grad(z.x, v.x), d1 = silu_grad(z.y, v.y);
                const float d2 = silu_grad(z.z, v.z), d3 = silu_grad(z.w, v.w);
                const uint32_t lo = (uint32_t)f2bf(d0) | ((uint32_t)f2bf(d1) << 16);
                const uint32_t hi = (uint32_t)f2bf(d2) | ((uint32_t)f2bf(d3) << 16);
                *reinterpret_cast<uint2*>(a.dzo + e) = uint2{lo, hi};
                cs.x += d0;
                cs.y += d1;
                cs.z += d2;
                cs.w += d3;
            }
        }
    }
}

template <int TN>
__device__ __forceinline__ void dgrad_silu_sums(const ConvArgs& a, float* red, int n0, int tile_m, const float4& cs) {
    constexpr int CG = TN / 4;
    static_assert(256 % CG == 0, "a thread's column group must stay fixed");
    const int t = (int)threadIdx.x;
    *reinterpret_cast<float4*>(&red[4 * t]) = cs;
    __syncthreads();
    for (int c = t; c < TN; c += 256) {
        const int n = n0 + c;
        if (n >= a.N) continue;
        float v = 0.f;
        for (int k = 0; k < 256 / CG; ++k) v += red[4 * ((c >> 2) + CG * k) + (c & 3)];
        a.bpart[(int64_t)tile_m * a.N + n] = v;
    }
}

// Index k = tap * C + c of a run along a (tap, channel) axis, advanced tile
// by tile without division.
struct TapChan {
    int tap, c;
    __device__ __forceinline__ void init(int k, int C) {
        tap = k / C;
        c = k - tap * C;
    }
    __device__ __forceinline__ void step(int dk, int C) {
        c += dk;
        while (c >= C) {
            c -= C;
            ++tap;
        }
    }
};

// A pixel index k = (b H + y) W + x, advanced tile by tile without division.
struct PixPos {
    int b, y, x;
    __device__ __forceinline__ void init(int k, int H, int W) {
        const int hw = H * W;
        b = k / hw;
        const int r = k - b * hw;
        y = r / W;
        x = r - y * W;
    }
    __device__ __forceinline__ void step(int dk, int H, int W) {
        x += dk;
        while (x >= W) {
            x -= W;
            if (++y >= H) {
                y = 0;
                ++b;
            }
        }
    }
};

// The 16-byte run at channel c of pixel (b, y + dy, x + dx) of a [B][H][W][C]
// tensor, zero outside the image.
__device__ __forceinline__ uint4 pix_run(const uint16_t* t, const ConvArgs& a, int C, int b, int y, int x, int c) {
    if (y < 0 || y >= a.H || x < 0 || x >= a.W) return zero_run();
    return *reinterpret_cast<const uint4*>(t + ((int64_t)(b * a.H + y) * a.W + x) * C + c);
}

template <int MODE, int TM, int TN, int WM, int WN>
__global__ void __launch_bounds__(256) conv_gemm2(ConvArgs a) {
    static_assert(WM * WN == 4 && TM % (16 * WM) == 0 && TN % (16 * WN) == 0, "tile shape");
    constexpr int FM = TM / WM / 16, FN = TN / WN / 16;  // 16 x 16 blocks per wave
    constexpr int RA = (TM * 4 + 255) / 256, RB = (TN * 4 + 255) / 256;  // staging runs per thread
    constexpr bool AR = a_rows_fast<MODE>(), BR = b_rows_fast<MODE>();
    // The staging buffers, and after the k loop the output slab of the
    // epilogue (below), in one LDS array.
    constexpr int A_ELEMS = tile_elems<TM, AR>(), B_ELEMS = tile_elems<TN, BR>();
    constexpr int SP = TN + 4;        // output slab row pitch (floats)
    constexpr int SR = TM / WM;       // output slab rows: one row of waves
    static_assert(SR * SP * 4 <= (2 * A_ELEMS + 2 * B_ELEMS) * 2, "output slab exceeds the staging LDS");
    __shared__ __attribute__((aligned(16))) uint16_t smem[2 * A_ELEMS + 2 * B_ELEMS];
    uint16_t(*As)[A_ELEMS] = reinterpret_cast<uint16_t(*)[A_ELEMS]>(smem);
    uint16_t(*Bs)[B_ELEMS] = reinterpret_cast<uint16_t(*)[B_ELEMS]>(smem + 2 * A_ELEMS);
    const int t = (int)threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave / WN, wn = wave % WN;
    // Pixel tiles (forward, data gradient): workgroups go to the 8 XCDs
    // round-robin (by linear index; gridDim.x tiles per row of the grid), so
    // XCD x is given the x-th eighth of the tiles in order -- the
    // neighbouring tiles whose tap gathers share image rows then meet in the
    // same L2.  Eighths of q or q + 1 tiles (the first r = nx mod 8 longer): a
    // bijection of 0 .. nx - 1.
    int bx = (int)blockIdx.x;
    if (MODE != MODE_WGRAD) {
        const int nx = (int)gridDim.x, q = nx >> 3, r = nx & 7, x = bx & 7;
        bx = x * q + min(x, r) + (bx >> 3);
    }
    const int m0 = bx * TM, n0 = (int)blockIdx.y * TN;
    int k_begin = 0, k_end = a.K;
    if (MODE == MODE_WGRAD) {
        k_begin = (int)blockIdx.z * a.kchunk;
        k_end = min(a.K, k_begin + a.kchunk);
    }
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Staging runs: run r of a T x 32 tile lies along k at row r / 4,
    // columns 8 (r % 4) .. + 7, or along rows (rows 8 (r % (T / 8)) .. + 7,
    // column r / (T / 8)); its row and column are the same in every k-tile,
    // so the gather positions are set up once and advanced by 32 per tile.
    int a_row[RA], a_kc[RA], b_row[RB], b_kc[RB];
    bool a_on[RA], b_on[RB];
    PixPos a_pix[RA], b_pix[RB];
    TapChan a_tc[RA], b_tc[RB];
#pragma unroll
    for (int q = 0; q < RA; ++q) {
        const int r = t + 256 * q;
        a_on[q] = r < TM * 4;
        a_row[q] = AR ? m0 + 8 * (r % (TM / 8)) : m0 + (r >> 2);
        a_kc[q] = AR ? r / (TM / 8) : 8 * (r & 3);
        a_on[q] = a_on[q] && a_row[q] < a.M;
        if (MODE == MODE_FWD || MODE == MODE_DGRAD) {
            // A = input (forward) / dz (data gradient) at pixel row, shifted by the tap of k.
            a_pix[q].init(a_on[q] ? a_row[q] : 0, a.H, a.W);
            a_tc[q].init(k_begin + a_kc[q], MODE == MODE_FWD ? a.Cin : a.Cout);
        }
    }
#pragma unroll
    for (int q = 0; q < RB; ++q) {
        const int r = t + 256 * q;
        b_on[q] = r < TN * 4;
        b_row[q] = BR ? n0 + 8 * (r % (TN / 8)) : n0 + (r >> 2);
        b_kc[q] = BR ? r / (TN / 8) : 8 * (r & 3);
        b_on[q] = b_on[q] && b_row[q] < a.N;
        if (MODE == MODE_DGRAD) b_tc[q].init(k_begin + b_kc[q], a.Cout);  // k = (tap, co)
        if (MODE == MODE_WGRAD) {
            b_tc[q].init(b_on[q] ? b_row[q] : 0, a.Cin);  // row = (tap, ci), fixed
            b_pix[q].init(k_begin + b_kc[q], a.H, a.W);    // k = pixel
        }
    }
    uint4 ra[RA], rb[RB];
    // Loads of the k-tile at k0 (called for k0 = k_begin, k_begin + 32, ...
    // in order: each call advances the runs' positions to the next tile).
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int k = k0 + a_kc[q];
            uint4 v = zero_run();
            if (a_on[q] && k < k_end) {
                if (MODE == MODE_FWD) {
                    const int tap = a_tc[q].tap, dy = tap / 3 - 1, dx = tap - 3 * (tap / 3) - 1;
                    v = pix_run(a.x, a, a.Cin, a_pix[q].b, a_pix[q].y + dy, a_pix[q].x + dx, a_tc[q].c);
                } else if (MODE == MODE_DGRAD) {
                    const int tap = a_tc[q].tap, dy = 1 - tap / 3, dx = 1 - (tap - 3 * (tap / 3));
                    v = pix_run(a.dz, a, a.Cout, a_pix[q].b, a_pix[q].y + dy, a_pix[q].x + dx, a_tc[q].c);
                } else {
                    v = *reinterpret_cast<const uint4*>(a.dz + (int64_t)k * a.Cout + a_row[q]);
                }
            }
            ra[q] = v;
            if (MODE != MODE_WGRAD) a_tc[q].step(BK2, MODE == MODE_FWD ? a.Cin : a.Cout);
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int k = k0 + b_kc[q];
            uint4 v = zero_run();
            if (b_on[q] && k < k_end) {
                if (MODE == MODE_FWD) {
                    v = *reinterpret_cast<const uint4*>(a.w + (int64_t)b_row[q] * a.K + k);
                } else if (MODE == MODE_DGRAD) {
                    v = *reinterpret_cast<const uint4*>(a.w + ((int64_t)b_tc[q].c * 9 + b_tc[q].tap) * a.Cin +
                                                        b_row[q]);
                } else {
                    const int tap = b_tc[q].tap, dy = tap / 3 - 1, dx = tap - 3 * (tap / 3) - 1;
                    v = pix_run(a.x, a, a.Cin, b_pix[q].b, b_pix[q].y + dy, b_pix[q].x + dx, b_tc[q].c);
                }
            }
            rb[q] = v;
            if (MODE == MODE_DGRAD) b_tc[q].step(BK2, a.Cout);
            if (MODE == MODE_WGRAD) b_pix[q].step(BK2, a.H, a.W);
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int r = t + 256 * q;
            if (r < TM * 4) stage_store2<TM, AR>(r, As[buf], ra[q]);
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int r = t + 256 * q;
            if (r < TN * 4) stage_store2<TN, BR>(r, Bs[buf], rb[q]);
        }
    };
    const int row = lane & 15;
    int buf = 0;
    if (k_begin < k_end) {
        fetch(k_begin);
        put(0);
    }
    __syncthreads();
    for (int k0 = k_begin; k0 < k_end; k0 += BK2) {
        const bool more = k0 + BK2 < k_end;
        if (more) fetch(k0 + BK2);
        bf16x8v fa[FM], fb[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) fa[i] = frag2<TM, AR>(As[buf], (wm * FM + i) * 16, lane);
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = frag2<TN, BR>(Bs[buf], (wn * FN + j) * 16, lane);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        if (more) put(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    float* out = a.out;
    if (MODE == MODE_WGRAD) out += (int64_t)blockIdx.z * a.M * a.N;
    const int kr = (lane >> 4) * 4;
    if constexpr (MODE == MODE_DGRAD && 256 % (TN / 4) == 0) {
        if (a.zact) {
            float* slab = reinterpret_cast<float*>(smem);
            float4 cs = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int h = 0; h < WM; ++h) {
                if (wm == h) {
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
#pragma unroll
                            for (int jj = 0; jj < 4; ++jj)
                                slab[(i * 16 + kr + jj) * SP + (wn * FN + j) * 16 + row] = acc[i][j][jj];
                }
                __syncthreads();
                dgrad_silu_rows<WM, TM, TN>(a, slab, m0, n0, h, cs);
                __syncthreads();
            }
            dgrad_silu_sums<TN>(a, slab, n0, bx, cs);
            return;
        }
    }
    if (MODE != MODE_FWD) {
        // Gradients: straight from the MFMA layout (16-float row pieces).
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const int m = m0 + (wm * FM + i) * 16 + kr + jj;
                    const int n = n0 + (wn * FN + j) * 16 + row;
                    if (m < a.M && n < a.N) out[(int64_t)m * a.N + n] = acc[i][j][jj];
                }
        return;
    }
    // Forward epilogue, one row of waves at a time: its accumulators go to an
    // LDS slab of SR x TN floats (the staging buffers are free: the k loop
    // ended on a barrier), from which each thread writes 4-float runs of
    // output rows (whole 64- and 128-byte row pieces of z and their bf16
    // SiLU values instead of the MFMA layout's 16-value pieces), with the
    // bias (N is a multiple of 8, so a run is wholly inside or outside the
    // matrix).  Measured slower for the gradients (one FP32 output).
    float* slab = reinterpret_cast<float*>(smem);
    constexpr int RUNS = SR * TN / 4, RPT = (RUNS + 255) / 256;
#pragma unroll
    for (int h = 0; h < WM; ++h) {
        if (wm == h) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj)
                        slab[(i * 16 + kr + jj) * SP + (wn * FN + j) * 16 + row] = acc[i][j][jj];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = t + 256 * q;
            if (RUNS % 256 == 0 || r < RUNS) {
                const int sr = r / (TN / 4), c4 = 4 * (r % (TN / 4));
                const int m = m0 + h * SR + sr, n = n0 + c4;
                if (m < a.M && n < a.N) {
                    float4 v = *reinterpret_cast<const float4*>(&slab[sr * SP + c4]);
                    if (a.bias) {
                        const float4 b = *reinterpret_cast<const float4*>(a.bias + n);
                        v.x += b.x;
                        v.y += b.y;
                        v.z += b.z;
                        v.w += b.w;
                    }
                    *reinterpret_cast<float4*>(out + (int64_t)m * a.N + n) = v;
                    if (a.act) {
                        const uint32_t lo = (uint32_t)silu_bf16(v.x) | ((uint32_t)silu_bf16(v.y) << 16);
                        const uint32_t hi = (uint32_t)silu_bf16(v.z) | ((uint32_t)silu_bf16(v.w) << 16);
                        *reinterpret_cast<uint2*>(a.act + (int64_t)m * a.N + n) = uint2{lo, hi};
                    }
                }
            }
        }
        __syncthreads();
    }
}

// The forward and data-gradient GEMMs with the three horizontal taps taken
// from one staged slab (the gathered operand's channel count a multiple of
// 32, N tiles of 32 or 64): stage s = (kernel row ky, 32-channel chunk cc)
// stages the TM + 2 consecutive pixels m0 - 1 + j + sdy W (j < TM + 2; sdy
// the row shift of ky) at the chunk's 32 channels, and the B tiles of the
// row's three taps; tap kx then reads the A fragment of output pixel m at
// staged pixel m - m0 + 1 + sdx (sdx its column shift).  A staged pixel is
// the tap's source pixel exactly when that source lies in the image; a lane
// whose source does not (the image's edge columns and rows, pixels past M)
// takes a zero fragment instead.  The gathered operand is read about 3 times
// per output tile instead of 9.  Forward: sdy = ky - 1, sdx = kx - 1, B(n,
// k) = w[n][tap][ci]; data gradient: sdy = 1 - ky, sdx = 1 - kx, B(n = ci,
// k = (tap, co)) = w[co][tap][ci] staged in source order.  Same pipeline as
// conv_gemm2 (two LDS stages, the next stage's runs loaded into registers
// while the current one is computed), same epilogues.
template <int MODE, int TM, int TN, int WM, int WN>
__global__ void __launch_bounds__(256) conv_gemm3(ConvArgs a) {
    static_assert(MODE == MODE_FWD || MODE == MODE_DGRAD, "pixel-row GEMMs only");
    static_assert(WM * WN == 4 && TM % (16 * WM) == 0 && TN % (16 * WN) == 0 && TN <= 64, "tile shape");
    constexpr int FM = TM / WM / 16, FN = TN / WN / 16;
    constexpr bool BR = MODE == MODE_DGRAD;
    constexpr int AROWS = TM + 2, A_ELEMS = AROWS * LDK2, B_ELEMS = tile_elems<TN, BR>();
    constexpr int STAGE = A_ELEMS + 3 * B_ELEMS;
    constexpr int RA = (AROWS * 4 + 255) / 256, RB = (3 * TN * 4 + 255) / 256;
    constexpr int SP = TN + 4, SR = TM / WM;
    static_assert(SR * SP * 4 <= 2 * STAGE * 2, "output slab exceeds the staging LDS");
    __shared__ __attribute__((aligned(16))) uint16_t smem[2 * STAGE];
    const int t = (int)threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int s16 = lane & 15, g = lane >> 4;
    int bx = (int)blockIdx.x;
    {
        const int nx = (int)gridDim.x, q = nx >> 3, r = nx & 7, x = bx & 7;
        bx = x * q + min(x, r) + (bx >> 3);
    }
    const int m0 = bx * TM, n0 = (int)blockIdx.y * TN;
    const int Cg = MODE == MODE_FWD ? a.Cin : a.Cout;  // channels of the gathered operand
    const uint16_t* src = MODE == MODE_FWD ? a.x : a.dz;
    const int ncc = Cg >> 5, nstage = 3 * ncc;
    // This lane's output pixels (one per M fragment): image row and column,
    // or y = -(1 << 20) past M (every tap then invalid).
    int py[FM], px[FM];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int m = m0 + (wm * FM + i) * 16 + s16;
        if (m < a.M) {
            const int r = m % (a.H * a.W);
            py[i] = r / a.W;
            px[i] = r - py[i] * a.W;
        } else {
            py[i] = -(1 << 20);
            px[i] = 0;
        }
    }
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    uint4 ra[RA], rb[RB];
    auto fetch = [&](int st) {
        const int ky = st / ncc, cc = st - ky * ncc;
        const int sdy = MODE == MODE_FWD ? ky - 1 : 1 - ky;
        const int64_t f0 = (int64_t)m0 - 1 + (int64_t)sdy * a.W;
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int r = t + 256 * q;
            const int64_t f = f0 + (r >> 2);
            ra[q] = (r < AROWS * 4 && f >= 0 && f < a.M)
                        ? *reinterpret_cast<const uint4*>(src + f * Cg + 32 * cc + 8 * (r & 3))
                        : zero_run();
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int r = t + 256 * q;
            uint4 v = zero_run();
            if (r < 3 * TN * 4) {
                const int kx = r / (TN * 4), rr = r - kx * (TN * 4), tap = 3 * ky + kx;
                if (MODE == MODE_FWD) {
                    const int n = n0 + (rr >> 2);
                    if (n < a.N) v = *reinterpret_cast<const uint4*>(a.w + (int64_t)n * a.K + tap * a.Cin + 32 * cc +
                                                                     8 * (rr & 3));
                } else {
                    const int co = 32 * cc + rr / (TN / 8), n = n0 + 8 * (rr % (TN / 8));
                    if (n < a.N) v = *reinterpret_cast<const uint4*>(a.w + ((int64_t)co * 9 + tap) * a.Cin + n);
                }
            }
            rb[q] = v;
        }
    };
    auto put = [&](int buf) {
        uint16_t* A = smem + buf * STAGE;
        uint16_t* Bb = A + A_ELEMS;
#pragma unroll
        for (int q = 0; q < RA; ++q) {
            const int r = t + 256 * q;
            if (r < AROWS * 4) *reinterpret_cast<uint4*>(&A[(r >> 2) * LDK2 + 8 * (r & 3)]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int r = t + 256 * q;
            if (r < 3 * TN * 4) {
                const int kx = r / (TN * 4), rr = r - kx * (TN * 4);
                stage_store2<TN, BR>(rr, Bb + kx * B_ELEMS, rb[q]);
            }
        }
    };
    auto mma = [&](int st, int buf) {
        const int ky = st / ncc;
        const int sdy = MODE == MODE_FWD ? ky - 1 : 1 - ky;
        const uint16_t* A = smem + buf * STAGE;
        const uint16_t* Bb = A + A_ELEMS;
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            const int sdx = MODE == MODE_FWD ? kx - 1 : 1 - kx;
            bf16x8v fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int y = py[i] + sdy, x = px[i] + sdx;
                const bool ok = y >= 0 && y < a.H && x >= 0 && x < a.W;
                const int row = (wm * FM + i) * 16 + s16 + 1 + sdx;
                const uint4 v = *reinterpret_cast<const uint4*>(&A[row * LDK2 + 8 * g]);
                fa[i] = __builtin_bit_cast(bf16x8v, ok ? v : zero_run());
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) fb[j] = frag2<TN, BR>(Bb + kx * B_ELEMS, (wn * FN + j) * 16, lane);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
    };
    fetch(0);
    put(0);
    __syncthreads();
    int buf = 0;
    for (int st = 0; st < nstage; ++st) {
        const bool more = st + 1 < nstage;
        if (more) fetch(st + 1);
        mma(st, buf);
        if (more) put(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    const int kr = g * 4;
    if constexpr (MODE == MODE_DGRAD) {
        if (a.zact) {
            float* slab = reinterpret_cast<float*>(smem);
            float4 cs = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int h = 0; h < WM; ++h) {
                if (wm == h) {
#pragma unroll
                    for (int i = 0; i < FM; ++i)
#pragma unroll
                        for (int j = 0; j < FN; ++j)
#pragma unroll
                            for (int jj = 0; jj < 4; ++jj)
                                slab[(i * 16 + kr + jj) * SP + (wn * FN + j) * 16 + s16] = acc[i][j][jj];
                }
                __syncthreads();
                dgrad_silu_rows<WM, TM, TN>(a, slab, m0, n0, h, cs);
                __syncthreads();
            }
            dgrad_silu_sums<TN>(a, slab, n0, bx, cs);
            return;
        }
    }
    if (MODE == MODE_DGRAD) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const int m = m0 + (wm * FM + i) * 16 + kr + jj;
                    const int n = n0 + (wn * FN + j) * 16 + s16;
                    if (m < a.M && n < a.N) a.out[(int64_t)m * a.N + n] = acc[i][j][jj];
                }
        return;
    }
    // Forward epilogue through an LDS slab, as in conv_gemm2.
    float* slab = reinterpret_cast<float*>(smem);
    constexpr int RUNS = SR * TN / 4, RPT = (RUNS + 255) / 256;
#pragma unroll
    for (int h = 0; h < WM; ++h) {
        if (wm == h) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj)
                        slab[(i * 16 + kr + jj) * SP + (wn * FN + j) * 16 + s16] = acc[i][j][jj];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = t + 256 * q;
            if (RUNS % 256 == 0 || r < RUNS) {
                const int sr = r / (TN / 4), c4 = 4 * (r % (TN / 4));
                const int m = m0 + h * SR + sr, n = n0 + c4;
                if (m < a.M && n < a.N) {
                    float4 v = *reinterpret_cast<const float4*>(&slab[sr * SP + c4]);
                    if (a.bias) {
                        const float4 b = *reinterpret_cast<const float4*>(a.bias + n);
                        v.x += b.x;
                        v.y += b.y;
                        v.z += b.z;
                        v.w += b.w;
                    }
                    *reinterpret_cast<float4*>(a.out + (int64_t)m * a.N + n) = v;
                    if (a.act) {
                        const uint32_t lo = (uint32_t)silu_bf16(v.x) | ((uint32_t)silu_bf16(v.y) << 16);
                        const uint32_t hi = (uint32_t)silu_bf16(v.z) | ((uint32_t)silu_bf16(v.w) << 16);
                        *reinterpret_cast<uint2*>(a.act + (int64_t)m * a.N + n) = uint2{lo, hi};
                    }
                }
            }
        }
        __syncthreads();
    }
}

// The weight gradient from LDS-resident row blocks, for channel counts that
// are multiples of 16: dw[co][tap][ci] = sum over pixels p of dz[p][co]
// x[p + shift(tap)][ci].  The pixel sum is the GEMM's k, against the NHWC
// layout of both operands.  Workgroup (co tile, ci tile, slice) takes rows
// y0 .. y1 - 1 of one image in chunks of RC rows: the chunk's dz rows (one
// LDS row per pixel, CO_T channels) and its x rows with a halo row and
// column on every side (zero outside the image, CI_T = 16 channels) are
// staged with 16-byte stores, and gfx950's LDS transpose read
// (ds_read_b64_tr_b16) hands the matrix cores each operand as 8 consecutive
// pixels of one channel: in a 16-lane group, lane s reads pixel k0 + s / 4
// at channels 4 (s % 4) .. + 3, and output lane 4 a + b receives channel b
// of the reads of lanes a, a + 4, a + 8, a + 12 -- pixels k0 .. k0 + 3 of
// channel 4 a + b (lane layout measured on the GPU).  Two such reads make the
// 8 k values of v_mfma_f32_16x16x32_bf16's operand.  A tap's shift is an
// address offset into the staged x rows, so the 9 taps read one tile.  Wave
// w accumulates every co fragment against taps w, w + 4, w + 8; the slice's
// sums go to out[slice] for the fixed-order column sum (deterministic).
// Pixels of a chunk past its last row (the k padding to a multiple of 32)
// have dz = 0 and read the chunk's last pixel's x.
struct WgArgs {
    const uint16_t* x;   // [B][H][W][Cin]
    const uint16_t* dz;  // [B][H][W][Cout]
    float* out;          // partial sums [S][Cout][9][Cin]
    int B, H, W, Cin, Cout;
    int RC;   // rows per chunk
    int rps;  // rows per slice (a multiple of RC)
    int spi;  // slices per image
};

constexpr int WG_CI = 16, WG_XP = WG_CI + 8;  // ci per workgroup; x tile row pitch (bf16, 48 bytes)
constexpr int WG_RUNS = 8;                     // staging runs per thread per tile, at most
__host__ __device__ constexpr int wg_dp(int co_t) { return co_t + 8; }  // dz tile row pitch

template <int CO_T>
__global__ void __launch_bounds__(256, 2) wgrad_lds(WgArgs a) {
    constexpr int DP = wg_dp(CO_T), FC = CO_T / 16, NW = 3;  // taps per wave: w, w + 4, w + 8 (< 9)
    extern __shared__ __attribute__((aligned(16))) uint16_t lds[];
    const int W = a.W, W2 = W + 2, RC = a.RC;
    const int KQ = (RC * W + 31) / 32 * 32;
    uint16_t* xs = lds;
    uint16_t* ds = lds + (RC + 2) * W2 * WG_XP;
    // Workgroups go to the 8 XCDs round-robin by linear index; the tiles are
    // handed out so that XCD x takes the x-th eighth of the (slice, tile)
    // list in order: the tiles of one slice, which stage the same dz and x
    // rows, then meet in one L2.
    const int nco = (int)gridDim.x, nci = (int)gridDim.y;
    const int n = nco * nci * (int)gridDim.z;
    const int lin = (int)blockIdx.x + nco * ((int)blockIdx.y + nci * (int)blockIdx.z);
    const int lq = n >> 3, lr = n & 7, lx = lin & 7;
    const int lg = lx * lq + min(lx, lr) + (lin >> 3);
    const int co0 = (lg % nco) * CO_T, ci0 = ((lg / nco) % nci) * WG_CI;
    const int sl = lg / (nco * nci), b = sl / a.spi;
    const int y0 = (sl - b * a.spi) * a.rps, y1 = min(a.H, y0 + a.rps);
    const int t = (int)threadIdx.x, lane = t & 63, wave = t >> 6;
    const int s = lane & 15, g = lane >> 4;
    const int qlane = 8 * g + (s >> 2);  // this lane's pixel in a k-step (half 0; +4 for half 1)
    const int cl = 4 * (s & 3);          // its channel offset within a fragment
    f32x4 acc[FC][NW];
#pragma unroll
    for (int i = 0; i < FC; ++i)
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t img = (int64_t)b * a.H;
    // Staging runs (16 bytes) of one chunk: x rows yc - 1 .. yc + nr,
    // columns -1 .. W, two 8-channel runs each; dz pixels 0 .. KQ - 1 (zero
    // past the chunk's last row).  Thread t moves runs t + 256 i (the plan
    // keeps both tiles within 256 WG_RUNS runs).  The next chunk's runs are
    // loaded into registers while the current chunk is computed.
    constexpr int RD = CO_T / 8;
    uint4 rx[WG_RUNS], rd[WG_RUNS];
    auto fetch = [&](int yc) {
        const int nr = min(RC, y1 - yc), np = nr * W, nx = (nr + 2) * W2 * 2;
#pragma unroll
        for (int i = 0; i < WG_RUNS; ++i) {
            const int r = t + 256 * i;
            const int pos = r >> 1, j = r & 1;
            const int row = pos / W2, col = pos - row * W2;
            const int y = yc - 1 + row, x = col - 1;
            rx[i] = (r < nx && y >= 0 && y < a.H && x >= 0 && x < W)
                        ? *reinterpret_cast<const uint4*>(a.x + ((img + y) * W + x) * a.Cin + ci0 + 8 * j)
                        : zero_run();
            const int q = r / RD, jd = r - q * RD;
            const int yy = q / W, xx = q - yy * W;
            rd[i] = (q < np) ? *reinterpret_cast<const uint4*>(a.dz + ((img + yc + yy) * W + xx) * a.Cout + co0 +
                                                               8 * jd)
                             : zero_run();
        }
    };
    auto put = [&](int yc) {
        const int nr = min(RC, y1 - yc), nx = (nr + 2) * W2 * 2;
#pragma unroll
        for (int i = 0; i < WG_RUNS; ++i) {
            const int r = t + 256 * i;
            if (r < nx) *reinterpret_cast<uint4*>(xs + (r >> 1) * WG_XP + 8 * (r & 1)) = rx[i];
            const int q = r / RD;
            if (q < KQ) *reinterpret_cast<uint4*>(ds + q * DP + 8 * (r - q * RD)) = rd[i];
        }
    };
    fetch(y0);
    for (int yc = y0; yc < y1; yc += RC) {
        const int nr = min(RC, y1 - yc), np = nr * W;
        put(yc);
        __syncthreads();
        if (yc + RC < y1) fetch(yc + RC);
        // This lane's pixels in the k-step (two halves) as (row, column) of
        // the chunk, advanced by 32 pixels per step.
        int py[2], px[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = min(qlane + 4 * h, np - 1);
            py[h] = q / W;
            px[h] = q - py[h] * W;
        }
        for (int kb = 0; kb < KQ; kb += 32) {
            bf16x8v fa[FC], fb[NW];
#pragma unroll
            for (int i = 0; i < FC; ++i) {
                const uint16_t* p0 = ds + (kb + qlane) * DP + 16 * i + cl;
                fa[i] = join8(tr_read(p0), tr_read(p0 + 4 * DP));
            }
#pragma unroll
            for (int j = 0; j < NW; ++j) {
                const int tap = wave + 4 * j;
                if (tap < 9) {
                    const int ky = tap / 3, kx = tap - 3 * ky;
                    const uint16_t* p0 = xs + ((py[0] + ky) * W2 + px[0] + kx) * WG_XP + cl;
                    const uint16_t* p1 = xs + ((py[1] + ky) * W2 + px[1] + kx) * WG_XP + cl;
                    fb[j] = join8(tr_read(p0), tr_read(p1));
                }
            }
#pragma unroll
            for (int i = 0; i < FC; ++i)
#pragma unroll
                for (int j = 0; j < NW; ++j)
                    if (wave + 4 * j < 9)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            // Next k-step: 32 pixels on, clamped to the chunk's last pixel.
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = min(kb + 32 + qlane + 4 * h, np - 1);
                py[h] = q / W;
                px[h] = q - py[h] * W;
            }
        }
        __syncthreads();
    }
    float* out = a.out + (int64_t)sl * a.Cout * 9 * a.Cin;
#pragma unroll
    for (int i = 0; i < FC; ++i)
#pragma unroll
        for (int j = 0; j < NW; ++j) {
            const int tap = wave + 4 * j;
            if (tap >= 9) continue;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int co = co0 + 16 * i + 4 * g + e, ci = ci0 + s;
                out[((int64_t)co * 9 + tap) * a.Cin + ci] = acc[i][j][e];
            }
        }
}

// out[i] = sum over rows s < S of part[s][i], i < n: a block of 16 waves
// per 64 outputs.  With n < 64 (one block) a wave's lanes take CW = the
// power of two >= n columns and 64 / CW interleaved row groups (lane l: column
// l % CW, rows l / CW, + 64 / CW, ...); wave w's rows start at w 64 / CW.
// Each lane keeps 8 running sums (eight independent loads in flight), added
// in a fixed tree; the row groups are combined by fixed xor shuffles and the
// 16 waves in a fixed tree order (deterministic, whatever the schedule).
__global__ void __launch_bounds__(1024) column_sum(const float* part, float* out, int64_t n, int S) {
    int cw = 64;
    if (n < 64) {
        cw = 1;
        while (cw < n) cw <<= 1;
    }
    const int R = 64 / cw;  // row groups per wave
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int64_t i = (int64_t)blockIdx.x * 64 + (lane % cw);
    const int step = 16 * R;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (i < n) {
        int z = w * R + lane / cw;
        for (; z + 7 * step < S; z += 8 * step)
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += part[(int64_t)(z + step * j) * n + i];
        for (int j = 0; z < S; z += step, ++j) acc[j] += part[(int64_t)z * n + i];
    }
    float s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    for (int off = cw; off < 64; off <<= 1) s += __shfl_xor(s, off);
    __shared__ float sh[16][64];
    sh[w][lane] = s;
    __syncthreads();
    for (int h = 8; h > 0; h >>= 1) {
        if (w < h) sh[w][lane] += sh[w + h][lane];
        __syncthreads();
    }
    if (w == 0 && lane < cw && i < n) out[i] = sh[0][lane];
}

// Per-channel pixel sums (the bias gradient), pass 1: block (cb, s) sums
// channels cb CW .. cb CW + CW - 1 over pixel slice s.  A wave row holds CW
// channels of 64 / CW consecutive pixels (CW = 64 when C is a multiple of
// 64, C when C divides 64, else 64 with the excess lanes idle), so every
// wave reads 64 consecutive values per step; partial sum q (pixels p0 + q,
// p0 + q + 4 PL, ...) of a channel sits at LDS slot q CW + channel and the
// 4 PL partials are added in a fixed tree order (deterministic).  For flat
// index i = p C + c, ld(i, p) loads the inputs and val(i, loaded) gives the
// value summed (it may also write per-element outputs); four pixels' loads
// are issued before their values are formed.
__host__ __device__ inline int bias_cw(int C) { return C % 64 == 0 ? 64 : (64 % C == 0 ? C : 64); }

template <class Ld, class Val>
__device__ __forceinline__ void channel_slice_sum(float* part, int P, int C, int rows, Ld ld, Val val) {
#pragma clang fp contract(off)  // each value rounded before it is added, as when it is read from memory
    const int CW = bias_cw(C), PL = 64 / CW;
    const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
    const int cc = lane % CW, q = w * PL + lane / CW;
    const int c = (int)blockIdx.x * CW + cc;
    const int p0 = (int)blockIdx.y * rows, p1 = min(P, p0 + rows);
    float s = 0.f;
    if (c < C) {
        // Four pixels per pass, their values summed in pixel order.
        const int step = 4 * PL;
        int p = p0 + q;
        for (; p + 3 * step < p1; p += 4 * step) {
            const int64_t i0 = (int64_t)p * C + c, i1 = i0 + (int64_t)step * C, i2 = i1 + (int64_t)step * C,
                          i3 = i2 + (int64_t)step * C;
            const auto t0 = ld(i0, p), t1 = ld(i1, p + step), t2 = ld(i2, p + 2 * step), t3 = ld(i3, p + 3 * step);
            s += val(i0, t0);
            s += val(i1, t1);
            s += val(i2, t2);
            s += val(i3, t3);
        }
        for (; p < p1; p += step) {
            const int64_t i = (int64_t)p * C + c;
            s += val(i, ld(i, p));
        }
    }
    __shared__ float sh[256];
    sh[t] = s;  // slot q CW + cc
    __syncthreads();
    for (int h = 2 * PL; h > 0; h >>= 1) {
        if (q < h) sh[q * CW + cc] += sh[(q + h) * CW + cc];
        __syncthreads();
    }
    if (q == 0 && c < C) part[(int64_t)blockIdx.y * C + c] = sh[cc];
}

__global__ void __launch_bounds__(256) bias_grad_part(const float* __restrict__ dz, float* __restrict__ part, int P,
                                                      int C, int rows) {
    channel_slice_sum(part, P, C, rows, [&](int64_t i, int) { return dz[i]; }, [](int64_t, float v) { return v; });
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__global__ void silu_fwd(const float* z, uint16_t* a, int64_t n) {
    GRID_STRIDE(i, n) a[i] = silu_bf16(z[i]);
}

// dz = da * silu'(z); written as bf16 (matrix-core operand) and FP32 (bias).
__global__ void silu_bwd(const float* z, const float* da, uint16_t* dzb, float* dzf, int64_t n) {
    GRID_STRIDE(i, n) {
        const float d = silu_grad(z[i], da[i]);
        if (dzf) dzf[i] = d;
        dzb[i] = f2bf(d);
    }
}

// The two fused: dz = da * silu'(z) (dz = da when z is null, a layer
// without activation) written as bf16, and its per-channel sums over pixel
// slice blockIdx.y into part, in channel_slice_sum's order.
template <bool ACT>
__global__ void __launch_bounds__(256) silu_bwd_bias_part(const float* __restrict__ z, const float* __restrict__ da,
                                                          uint16_t* __restrict__ dzb, float* __restrict__ part, int P,
                                                          int C, int rows) {
    channel_slice_sum(
        part, P, C, rows, [&](int64_t i, int) { return float2{ACT ? z[i] : 0.f, da[i]}; },
        [&](int64_t i, float2 t) {
            const float d = ACT ? silu_grad(t.x, t.y) : t.y;
            dzb[i] = f2bf(d);
            return d;
        });
}

// The encoder join's backward and the SiLU backward in one pass: da =
// 0.25 dy[b][Y/2][X/2][c] + skip[b][Y][X][c] (pool_bwd_add's value, the
// product rounded before the sum) is formed in registers and goes through
// silu_bwd_bias_part's dz and bias slices, so the FP32 da never reaches
// memory.  W is the fine width; B H W C < 2^31 and H, W even, so the pooled
// row of fine row r = b H + Y is r / 2.  r = p / W is (p wm) >> wsh
// (div_magic).
__global__ void __launch_bounds__(256) pool_silu_bwd_bias_part(const float* __restrict__ dy,
                                                               const float* __restrict__ skip,
                                                               const float* __restrict__ z, uint16_t* __restrict__ dzb,
                                                               float* __restrict__ part, int P, int C, int W,
                                                               uint64_t wm, int wsh, int rows) {
    channel_slice_sum(
        part, P, C, rows,
        [&](int64_t i, int p) {
            const int c = (int)i - p * C, r = (int)(((uint64_t)(uint32_t)p * wm) >> wsh), X = p - r * W;
            const int j = ((r >> 1) * (W >> 1) + (X >> 1)) * C + c;
            return float3{z[i], dy[j], skip[i]};
        },
        [&](int64_t i, float3 t) {
#pragma clang fp contract(off)
            const float d = silu_grad(t.x, 0.25f * t.y + t.z);
            dzb[i] = f2bf(d);
            return d;
        });
}

__global__ void to_bf16(const float* x, uint16_t* y, int64_t n) {
    GRID_STRIDE(i, n) y[i] = f2bf(x[i]);
}

__global__ void avgpool2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
    const int Ho = H / 2, Wo = W / 2;
    const int64_t n = (int64_t)B * Ho * Wo * C;
    GRID_STRIDE(i, n) {
        const int c = (int)(i % C);
        int64_t r = i / C;
        const int xo = (int)(r % Wo);
        r /= Wo;
        const int yo = (int)(r % Ho);
        const int b = (int)(r / Ho);
        const int64_t base = ((int64_t)(b * H + 2 * yo) * W + 2 * xo) * C + c;
        const float s = bf2f(x[base]) + bf2f(x[base + C]) + bf2f(x[base + (int64_t)W * C]) +
                        bf2f(x[base + (int64_t)W * C + C]);
        y[i] = f2bf(0.25f * s);
    }
}

__global__ void avgpool2_bwd(const float* dy, float* dx, int B, int H, int W, int C) {
    const int64_t n = (int64_t)B * H * W * C;
    const int Ho = H / 2, Wo = W / 2;
    GRID_STRIDE(i, n) {
        const int c = (int)(i % C);
        int64_t r = i / C;
        const int x = (int)(r % W);
        r /= W;
        const int y = (int)(r % H);
        const int b = (int)(r / H);
        dx[i] = 0.25f * dy[((int64_t)(b * Ho + y / 2) * Wo + x / 2) * C + c];
    }
}

// Nearest-neighbour x2: y[b][Y][X][c] = x[b][Y/2][X/2][c] for Y < 2H, X < 2W.
__global__ void upsample2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C) {
    const int Ho = 2 * H, Wo = 2 * W;
    const int64_t n = (int64_t)B * Ho * Wo * C;
    GRID_STRIDE(i, n) {
        const int c = (int)(i % C);
        int64_t r = i / C;
        const int X = (int)(r % Wo);
        r /= Wo;
        const int Y = (int)(r % Ho);
        const int b = (int)(r / Ho);
        y[i] = x[((int64_t)(b * H + Y / 2) * W + X / 2) * C + c];
    }
}

__global__ void upsample2_bwd(const float* dy, float* dx, int B, int H, int W, int C) {
    const int Wo = 2 * W;
    const int64_t n = (int64_t)B * H * W * C;
    GRID_STRIDE(i, n) {
        const int c = (int)(i % C);
        int64_t r = i / C;
        const int x = (int)(r % W);
        r /= W;
        const int y = (int)(r % H);
        const int b = (int)(r / H);
        const int64_t base = ((int64_t)(b * 2 * H + 2 * y) * Wo + 2 * x) * C + c;
        dx[i] = dy[base] + dy[base + C] + dy[base + (int64_t)Wo * C] + dy[base + (int64_t)Wo * C + C];
    }
}

// y[p] = [a[p] (Ca channels), b[p] (Cb channels)].
__global__ void concat_fwd(const uint16_t* a, int Ca, const uint16_t* b, int Cb, uint16_t* y, int64_t P) {
    const int C = Ca + Cb;
    const int64_t n = P * C;
    GRID_STRIDE(i, n) {
        const int64_t p = i / C;
        const int c = (int)(i - p * C);
        y[i] = (c < Ca) ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
    }
}

// da (+)= dy[:, :Ca], db (+)= dy[:, Ca:]; accumulate when acc_a / acc_b.
__global__ void concat_bwd(const float* dy, float* da, int Ca, int acc_a, float* db, int Cb, int acc_b, int64_t P) {
    const int C = Ca + Cb;
    const int64_t n = P * C;
    GRID_STRIDE(i, n) {
        const int64_t p = i / C;
        const int c = (int)(i - p * C);
        const float v = dy[i];
        if (c < Ca) {
            float* d = &da[p * Ca + c];
            *d = acc_a ? *d + v : v;
        } else {
            float* d = &db[p * Cb + (c - Ca)];
            *d = acc_b ? *d + v : v;
        }
    }
}

__global__ void add_inplace(float* y, const float* x, int64_t n) {
    GRID_STRIDE(i, n) y[i] += x[i];
}

// The decoder's input in one pass: y[b][Y][X] = [x[b][Y/2][X/2] (Cu
// channels), skip[b][Y][X] (Cs channels)] over the 2H x 2W fine grid, as
// upsample2_fwd then concat_fwd.  One thread per 16-byte run of y (channel
// counts multiples of 8), 32-bit indices (the host checks the run count).
__global__ void upcat_fwd(const uint16_t* x, int Cu, const uint16_t* skip, int Cs, uint16_t* y, int B, int H, int W) {
    const int Ho = 2 * H, Wo = 2 * W, R = (Cu + Cs) / 8;
    const int n = B * Ho * Wo * R;
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        const int p = i / R, j = i - p * R;
        const int X = p % Wo, r = p / Wo, Y = r % Ho, b = r / Ho;
        const int c = 8 * j;
        const uint4 v = (c < Cu) ? *reinterpret_cast<const uint4*>(x + ((int64_t)(b * H + Y / 2) * W + X / 2) * Cu + c)
                                 : *reinterpret_cast<const uint4*>(skip + (int64_t)p * Cs + (c - Cu));
        *reinterpret_cast<uint4*>(y + (int64_t)p * (Cu + Cs) + c) = v;
    }
}

// Its backward in one pass: from dy [B][2H][2W][Cu + Cs], dx[b][y][x][c] =
// the sum of dy over the 2 x 2 fine pixels of (y, x) at channel c < Cu (in
// upsample2_bwd's order) and dskip[fine pixel][c'] = dy[...][Cu + c'], as
// concat_bwd then upsample2_bwd.  One thread per coarse pixel and 4-float
// run (channel counts multiples of 4), 32-bit indices.
__global__ void upcat_bwd(const float* dy, float* dx, int Cu, float* dskip, int Cs, int B, int H, int W) {
    const int Wo = 2 * W, C = Cu + Cs, R = C / 4;
    const int n = B * H * W * R;
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        const int q = i / R, j = i - q * R;
        const int x = q % W, r = q / W, y = r % H, b = r / H;
        const int c = 4 * j;
        const int64_t f0 = (int64_t)(b * 2 * H + 2 * y) * Wo + 2 * x;  // fine pixels f0, f0 + 1, f0 + Wo, f0 + Wo + 1
        const int64_t f[4] = {f0, f0 + 1, f0 + Wo, f0 + Wo + 1};
        float4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = *reinterpret_cast<const float4*>(dy + f[k] * C + c);
        if (c < Cu) {
            float4 o;
            o.x = v[0].x + v[1].x + v[2].x + v[3].x;
            o.y = v[0].y + v[1].y + v[2].y + v[3].y;
            o.z = v[0].z + v[1].z + v[2].z + v[3].z;
            o.w = v[0].w + v[1].w + v[2].w + v[3].w;
            *reinterpret_cast<float4*>(dx + (int64_t)q * Cu + c) = o;
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) *reinterpret_cast<float4*>(dskip + f[k] * Cs + (c - Cu)) = v[k];
        }
    }
}

// dx = avgpool2 backward of dy (H, W the fine sizes) plus skip, as
// avgpool2_bwd then add_inplace: dx[b][Y][X][c] = 0.25 dy[b][Y/2][X/2][c] +
// skip[b][Y][X][c].  One thread per 4-float run (C a multiple of 4).
__global__ void pool_bwd_add(const float* dy, const float* skip, float* dx, int B, int H, int W, int C) {
#pragma clang fp contract(off)  // the product rounded before the sum, as in the separate passes
    const int Ho = H / 2, Wo = W / 2, R = C / 4;
    const int n = B * H * W * R;
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += (int)(gridDim.x * blockDim.x)) {
        const int p = i / R, j = i - p * R;
        const int X = p % W, r = p / W, Y = r % H, b = r / H;
        const float4 g = *reinterpret_cast<const float4*>(dy + ((int64_t)(b * Ho + Y / 2) * Wo + X / 2) * C + 4 * j);
        const float4 k = *reinterpret_cast<const float4*>(skip + (int64_t)p * C + 4 * j);
        float4 o;
        o.x = 0.25f * g.x + k.x;
        o.y = 0.25f * g.y + k.y;
        o.z = 0.25f * g.z + k.z;
        o.w = 0.25f * g.w + k.w;
        *reinterpret_cast<float4*>(dx + (int64_t)p * C + 4 * j) = o;
    }
}

// Mean squared error over n values: dy = 2 (y - t) / n; per-block partial
// sums of (y - t)^2 into part[blockIdx.x] (the host sums them).
__global__ void __launch_bounds__(256) mse(const float* y, const float* t, float* dy, float* part, int64_t n) {
    float s = 0.f;
    const float sc = 2.f / (float)n;
    GRID_STRIDE(i, n) {
        const float d = y[i] - t[i];
        s += d * d;
        if (dy) dy[i] = sc * d;
    }
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015) with bias correction on the gradient scaled by
// gscale (global-norm clipping); writes the bf16 copy.
__device__ __forceinline__ void adam_elem(float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t i,
                                          float lr, float b1, float b2, float eps, float c1, float c2, float gscale) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

__global__ void adam(float* w, uint16_t* wb, const float* g, float* m, float* v, int64_t n, float lr, float b1,
                     float b2, float eps, float c1, float c2, float gscale) {
    GRID_STRIDE(i, n) adam_elem(w, wb, g, m, v, i, lr, b1, b2, eps, c1, c2, gscale);
}

// Adam over up to kAdamMax tensors in one launch: tensor k owns blocks
// block0 .. block0 + ceil(n / 1024) - 1, 1024 elements each (4 per thread).
constexpr int kAdamMax = 32;
struct AdamSeg {
    float* w;
    uint16_t* wb;
    const float* g;
    float* m;
    float* v;
    int64_t n;
    int64_t block0;
};
struct AdamTable {
    AdamSeg s[kAdamMax];
    int count;
};

__global__ void __launch_bounds__(256) adam_multi(AdamTable t, float lr, float b1, float b2, float eps, float c1,
                                                  float c2, float gscale) {
    const int64_t b = blockIdx.x;
    int k = 0;
    while (k + 1 < t.count && t.s[k + 1].block0 <= b) ++k;
    const AdamSeg sg = t.s[k];
    const int64_t i0 = (b - sg.block0) * 1024 + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t i = i0 + 256 * j;
        if (i < sg.n) adam_elem(sg.w, sg.wb, sg.g, sg.m, sg.v, i, lr, b1, b2, eps, c1, c2, gscale);
    }
}

// Per-block partial sums of x[i]^2 (part[blockIdx.x]; fixed tree order).
__global__ void __launch_bounds__(256) sumsq(const float* x, float* part, int64_t n) {
    float s = 0.f;
    GRID_STRIDE(i, n) s += x[i] * x[i];
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dst[j] = src[idx[j]] (rows of row_len floats).
__global__ void gather_rows(const float* src, const int* idx, float* dst, int nrows, int64_t row_len) {
    const int64_t n = (int64_t)nrows * row_len;
    GRID_STRIDE(i, n) {
        const int64_t j = i / row_len;
        dst[i] = src[(int64_t)idx[j] * row_len + (i - j * row_len)];
    }
}

unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)); }

int finish(hipError_t e) {
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        g_nn_error = hipGetErrorString(e);
        return MPD_NN_E_DEVICE;
    }
    return MPD_NN_OK;
}

int bad(const char* msg) {
    g_nn_error = msg;
    return MPD_NN_E_ARG;
}

bool conv_dims_ok(int B, int H, int W, int Cin, int Cout) {
    return B > 0 && H > 0 && W > 0 && Cin > 0 && Cout > 0 && (int64_t)B * H * W < (1LL << 30) &&
           (int64_t)9 * Cin < (1 << 20) && (int64_t)9 * Cout < (1 << 20);
}

// Pixel slices of the bias gradient: about 128 pixels each, at most 1024.
int bias_slices(int P, int C) {
    (void)C;
    return std::max(1, std::min(1024, P / 128));
}

// Tile of the matrix-core GEMM for C = M x N (channel counts multiples of
// 8): 128 x N-wide tiles for the pixel-major passes, Cout-high tiles for the
// weight gradient.
struct Tile {
    int tm, tn;
};
Tile pick_tile(int M, int N) {
    Tile t;
    t.tn = N <= 16 ? 16 : N <= 32 ? 32 : N <= 64 ? 64 : 128;
    t.tm = M <= 16 ? 16 : M <= 32 ? 32 : M <= 64 ? 64 : 128;
    if (t.tm == 16 && t.tn < 64) t.tn = 64;  // four waves need 64 lanes of N then
    return t;
}
// The forward and data-gradient tiles: the widest tile, halved (rows
// first) down to 32 x 64 until about 1024 workgroups run, four per CU: the
// coarse levels' GEMMs (8400 pixels x 256 channels) would otherwise leave
// most CUs idle, one 128 x 128 tile per 64 k values of latency each.
Tile pick_tile_fill(int M, int N) {
    Tile t = pick_tile(M, N);
    // Multiples of 96 (the decoders' concatenated inputs, 96 and 192
    // channels) in 96-wide tiles: no idle quarter of a 128-wide one.
    if (N > 64 && N % 96 == 0 && N % 128 != 0) t.tn = 96;
    auto count = [&](const Tile& u) { return (long long)((M + u.tm - 1) / u.tm) * ((N + u.tn - 1) / u.tn); };
    // Rows halve while at least 32 and at least the tile's width (the
    // instantiated shapes below have 64 rows or more when 16 wide).
    auto rows_ok = [&](const Tile& u) { return u.tm > 32 && (u.tn >= 32 || u.tm > 64); };
    while (count(t) < 1024) {
        if (rows_ok(t) && t.tm >= t.tn) t.tm /= 2;
        else if (t.tn > 64) t.tn = 64;
        else if (rows_ok(t)) t.tm /= 2;
        else break;
    }
    return t;
}
bool vec_ok(int Cin, int Cout) { return Cin % 8 == 0 && Cout % 8 == 0; }

// Pixel splits of the weight gradient: enough workgroups to fill the GPU
// (about 2048 over the output tiles), each split at least 256 pixels.
int wgrad_splits(int P, int M, int N, bool vec) {
    int tm = BM, tn = BN;
    if (vec) {
        const Tile t = pick_tile(M, N);
        tm = t.tm;
        tn = t.tn;
    }
    const int tiles = ((M + tm - 1) / tm) * ((N + tn - 1) / tn);
    const int want = (2048 + tiles - 1) / tiles;
    return std::max(1, std::min(want, P / 256));
}
int wgrad_chunk(int P, int S) { return ((P + S - 1) / S + BK - 1) / BK * BK; }

// Launch plan of wgrad_lds: channel counts multiples of 16, rows per chunk
// for about 512 pixels with the LDS tiles within 60 KiB (two workgroups per
// CU), slices of whole chunks; false when a single row's tiles do not fit
// (very wide images: the GEMM path then).  The chunks per slice minimize
// rounds x chunks per slice, a round being the 512 workgroups MI355X holds
// at once (256 CUs, two each: the kernel's registers and LDS), over slice
// counts up to about 1024 workgroups (each slice adds a partial sum), ties
// to fewer slices: 800 workgroups of 4 chunks (two rounds, the second 56 %
// full) become 464 of 7 (one round).
struct WgPlan {
    int co_t, RC, rps, spi, S;
    size_t lds;
};
size_t wgrad_lds_bytes(int W, int RC, int co_t) {
    return ((size_t)(RC + 2) * (W + 2) * WG_XP + (size_t)((RC * W + 31) / 32 * 32) * wg_dp(co_t)) * 2;
}
bool wgrad_lds_plan(int B, int H, int W, int Cin, int Cout, WgPlan& p) {
    if (Cin % WG_CI != 0 || Cout % 16 != 0) return false;
    p.co_t = (Cout % 64 == 0) ? 64 : (Cout % 32 == 0) ? 32 : 16;
    constexpr size_t kMaxLds = 60 * 1024;
    int RC = std::max(1, std::min(H, 512 / W));
    // Both staged tiles within 256 WG_RUNS 16-byte runs (the kernel's
    // per-thread staging registers).
    auto fits = [&](int rc) {
        return wgrad_lds_bytes(W, rc, p.co_t) <= kMaxLds && (rc + 2) * (W + 2) * 2 <= 256 * WG_RUNS &&
               (rc * W + 31) / 32 * 32 * (p.co_t / 8) <= 256 * WG_RUNS;
    };
    while (RC > 1 && !fits(RC)) --RC;
    if (!fits(RC)) return false;
    p.RC = RC;
    p.lds = wgrad_lds_bytes(W, RC, p.co_t);
    const int tiles = (Cout / p.co_t) * (Cin / WG_CI);
    const int chunks = (H + RC - 1) / RC;
    constexpr long long kSlots = 512;
    // At most about 1024 workgroups: more slices cost more partial sums.
    const int want = std::max(1, 1024 / (tiles * B));
    const int cps_min = (chunks + std::min(chunks, want) - 1) / std::min(chunks, want);
    int cps = chunks;
    long long best = -1;
    for (int c = chunks; c >= cps_min; --c) {
        const long long wgs = (long long)tiles * B * ((chunks + c - 1) / c);
        const long long cost = (wgs + kSlots - 1) / kSlots * c;
        if (best < 0 || cost < best) {
            best = cost;
            cps = c;
        }
    }
    p.rps = cps * RC;
    p.spi = (H + p.rps - 1) / p.rps;
    p.S = B * p.spi;
    return true;
}

// Launch of conv_gemm2 for C = a.M x a.N with the tile pick_tile (weight
// gradient: its splits are sized to that tile) or pick_tile_fill chooses.
template <int MODE>
int launch_gemm2(const ConvArgs& a, int splits, hipStream_t s) {
    const Tile t = MODE == MODE_WGRAD ? pick_tile(a.M, a.N) : pick_tile_fill(a.M, a.N);
    const dim3 g((unsigned)((a.M + t.tm - 1) / t.tm), (unsigned)((a.N + t.tn - 1) / t.tn), (unsigned)splits);
    if constexpr (MODE != MODE_WGRAD) {
        // Gathered channels in 32-channel chunks: the three-tap slab kernel
        // for the tiles where it measured faster (32-wide tiles and 32-pixel
        // tiles; the 64 x 64 and 128 x 64 slab tiles hold three taps of B in
        // 41-52 KB of LDS, and the lower occupancy made them 4-5 % slower).
        if ((MODE == MODE_FWD ? a.Cin : a.Cout) % 32 == 0) {
#define MPD_GEMM3(TM, TN, WM, WN) \
    if (t.tm == TM && t.tn == TN) { hipLaunchKernelGGL((conv_gemm3<MODE, TM, TN, WM, WN>), g, dim3(256), 0, s, a); return (int)g.x; }
            MPD_GEMM3(128, 32, 4, 1)
            MPD_GEMM3(64, 32, 4, 1)
            MPD_GEMM3(32, 64, 2, 2)
            MPD_GEMM3(32, 32, 2, 2)
#undef MPD_GEMM3
        }
    }
#define MPD_GEMM2(TM, TN, WM, WN) \
    if (t.tm == TM && t.tn == TN) { hipLaunchKernelGGL((conv_gemm2<MODE, TM, TN, WM, WN>), g, dim3(256), 0, s, a); return (int)g.x; }
    MPD_GEMM2(128, 128, 2, 2)
    MPD_GEMM2(128, 96, 2, 2)
    MPD_GEMM2(128, 64, 2, 2)
    MPD_GEMM2(128, 32, 4, 1)
    MPD_GEMM2(128, 16, 4, 1)
    MPD_GEMM2(64, 128, 2, 2)
    MPD_GEMM2(64, 96, 2, 2)
    MPD_GEMM2(64, 64, 2, 2)
    MPD_GEMM2(64, 32, 4, 1)
    MPD_GEMM2(64, 16, 4, 1)
    MPD_GEMM2(32, 128, 2, 2)
    MPD_GEMM2(32, 96, 2, 2)
    MPD_GEMM2(32, 64, 2, 2)
    MPD_GEMM2(32, 32, 2, 2)
    MPD_GEMM2(16, 128, 1, 4)
    MPD_GEMM2(16, 64, 1, 4)
#undef MPD_GEMM2
    // Any other shape (every access is bounds-checked): the widest tile.
    const dim3 g2((unsigned)((a.M + 127) / 128), (unsigned)((a.N + 127) / 128), (unsigned)splits);
    hipLaunchKernelGGL((conv_gemm2<MODE, 128, 128, 2, 2>), g2, dim3(256), 0, s, a);
    return (int)g2.x;
}

// The forward pass, z = conv(x, w) + bias, with bf16(silu(z)) written to
// act as well when act is not null (in the GEMM's epilogue).
void launch_conv_fwd(const uint16_t* x, const uint16_t* w, const float* bias, float* z, uint16_t* act, int B, int H,
                     int W, int Cin, int Cout, hipStream_t s) {
    ConvArgs a{};
    a.x = x;
    a.w = w;
    a.bias = bias;
    a.out = z;
    a.act = act;
    a.B = B, a.H = H, a.W = W, a.Cin = Cin, a.Cout = Cout;
    a.M = B * H * W;
    a.N = Cout;
    a.K = 9 * Cin;
    const dim3 g((unsigned)((a.M + BM - 1) / BM), (unsigned)((a.N + BN - 1) / BN));
    if (vec_ok(Cin, Cout)) launch_gemm2<MODE_FWD>(a, 1, s);
    else hipLaunchKernelGGL((conv_gemm<MODE_FWD, false>), g, dim3(256), 0, s, a);
}

}  // namespace

// floor(n / d) = (n m) >> sh for 0 <= n < 2^31 and d >= 1, with l =
// ceil(log2 d), sh = 31 + l and m = ceil(2^sh / d) <= 2^32, so n m < 2^63
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, theorem 4.2).
struct DivMagic {
    uint64_t m;
    int sh;
};
static DivMagic div_magic(int d) {
    int l = 0;
    while ((1LL << l) < d) ++l;
    const int sh = 31 + l;
    return {((1ULL << sh) + (uint64_t)d - 1) / (uint64_t)d, sh};
}

extern "C" {

const char* mpd_nn_last_error(void) { return g_nn_error.c_str(); }

int mpd_nn_malloc(void** p, long long bytes) {
    if (!p || bytes < 0) return bad("malloc: bad argument");
    *p = nullptr;
    return finish(hipMalloc(p, (size_t)std::max(bytes, 1LL)));
}

int mpd_nn_free(void* p) { return finish(p ? hipFree(p) : hipSuccess); }

int mpd_nn_memcpy_h2d(void* dst, const void* src, long long bytes) {
    if (!dst || !src || bytes < 0) return bad("memcpy_h2d: bad argument");
    return finish(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
}

int mpd_nn_memcpy_d2h(void* dst, const void* src, long long bytes) {
    if (!dst || !src || bytes < 0) return bad("memcpy_d2h: bad argument");
    return finish(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
}

int mpd_nn_memset(void* p, int value, long long bytes) {
    if (!p || bytes < 0) return bad("memset: bad argument");
    return finish(hipMemset(p, value, (size_t)bytes));
}

int mpd_nn_synchronize(void) { return finish(hipDeviceSynchronize()); }

int mpd_nn_conv3x3_fwd(const uint16_t* x, const uint16_t* w, const float* bias, float* z, int B, int H, int W,
                       int Cin, int Cout, void* stream) {
    if (!x || !w || !z || !conv_dims_ok(B, H, W, Cin, Cout)) return bad("conv3x3_fwd: bad argument");
    launch_conv_fwd(x, w, bias, z, nullptr, B, H, W, Cin, Cout, (hipStream_t)stream);
    return finish(hipSuccess);
}

int mpd_nn_conv3x3_fwd_silu(const uint16_t* x, const uint16_t* w, const float* bias, float* z, uint16_t* a, int B,
                            int H, int W, int Cin, int Cout, void* stream) {
    if (!x || !w || !z || !a || !conv_dims_ok(B, H, W, Cin, Cout)) return bad("conv3x3_fwd_silu: bad argument");
    launch_conv_fwd(x, w, bias, z, a, B, H, W, Cin, Cout, (hipStream_t)stream);
    return finish(hipSuccess);
}

int mpd_nn_conv3x3_bwd_data(const uint16_t* dz, const uint16_t* w, float* dx, int B, int H, int W, int Cin, int Cout,
                            void* stream) {
    if (!dz || !w || !dx || !conv_dims_ok(B, H, W, Cin, Cout)) return bad("conv3x3_bwd_data: bad argument");
    ConvArgs a{};
    a.dz = dz;
    a.w = w;
    a.out = dx;
    a.B = B, a.H = H, a.W = W, a.Cin = Cin, a.Cout = Cout;
    a.M = B * H * W;
    a.N = Cin;
    a.K = 9 * Cout;
    const dim3 g((unsigned)((a.M + BM - 1) / BM), (unsigned)((a.N + BN - 1) / BN));
    if (vec_ok(Cin, Cout)) launch_gemm2<MODE_DGRAD>(a, 1, (hipStream_t)stream);
    else hipLaunchKernelGGL((conv_gemm<MODE_DGRAD, false>), g, dim3(256), 0, (hipStream_t)stream, a);
    return finish(hipSuccess);
}

long long mpd_nn_bwd_data_silu_work_floats(int B, int H, int W, int Cin, int Cout) {
    if (!conv_dims_ok(B, H, W, Cin, Cout)) return -1;
    return ((long long)B * H * W + 15) / 16 * Cin;  // pixel tiles of at least 16 rows
}

int mpd_nn_conv3x3_bwd_data_silu(const uint16_t* dz, const uint16_t* w, const float* z_in, uint16_t* dz_in_bf16,
                                 float* db_in, float* work, int B, int H, int W, int Cin, int Cout, void* stream) {
    if (!dz || !w || !z_in || !dz_in_bf16 || !db_in || !work || !conv_dims_ok(B, H, W, Cin, Cout) ||
        !vec_ok(Cin, Cout) || (Cin % 96 == 0 && Cin % 128 != 0))
        return bad("conv3x3_bwd_data_silu: bad argument (channel counts multiples of 8; Cin not 96 x odd)");
    ConvArgs a{};
    a.dz = dz;
    a.w = w;
    a.zact = z_in;
    a.dzo = dz_in_bf16;
    a.bpart = work;
    a.B = B, a.H = H, a.W = W, a.Cin = Cin, a.Cout = Cout;
    a.M = B * H * W;
    a.N = Cin;
    a.K = 9 * Cout;
    hipStream_t s = (hipStream_t)stream;
    const int tiles = launch_gemm2<MODE_DGRAD>(a, 1, s);
    hipLaunchKernelGGL(column_sum, dim3((unsigned)((Cin + 63) / 64)), dim3(1024), 0, s, (const float*)work, db_in,
                       (int64_t)Cin, tiles);
    return finish(hipSuccess);
}

long long mpd_nn_wgrad_work_floats(int B, int H, int W, int Cin, int Cout) {
    if (!conv_dims_ok(B, H, W, Cin, Cout)) return -1;
    WgPlan pl;
    if (wgrad_lds_plan(B, H, W, Cin, Cout, pl)) return (long long)pl.S * Cout * 9 * Cin;
    const int P = B * H * W;
    return (long long)wgrad_splits(P, Cout, 9 * Cin, vec_ok(Cin, Cout)) * Cout * 9 * Cin;
}

int mpd_nn_conv3x3_bwd_weight(const uint16_t* x, const uint16_t* dz, float* dw, float* work, int B, int H, int W,
                              int Cin, int Cout, void* stream) {
    if (!x || !dz || !dw || !work || !conv_dims_ok(B, H, W, Cin, Cout)) return bad("conv3x3_bwd_weight: bad argument");
    ConvArgs a{};
    a.x = x;
    a.dz = dz;
    a.out = work;
    a.B = B, a.H = H, a.W = W, a.Cin = Cin, a.Cout = Cout;
    hipStream_t s = (hipStream_t)stream;
    WgPlan pl;
    if (wgrad_lds_plan(B, H, W, Cin, Cout, pl)) {
        WgArgs w{x, dz, work, B, H, W, Cin, Cout, pl.RC, pl.rps, pl.spi};
        const dim3 g((unsigned)(Cout / pl.co_t), (unsigned)(Cin / WG_CI), (unsigned)pl.S);
        if (pl.co_t == 64) hipLaunchKernelGGL(wgrad_lds<64>, g, dim3(256), pl.lds, s, w);
        else if (pl.co_t == 32) hipLaunchKernelGGL(wgrad_lds<32>, g, dim3(256), pl.lds, s, w);
        else hipLaunchKernelGGL(wgrad_lds<16>, g, dim3(256), pl.lds, s, w);
        const int64_t n = (int64_t)Cout * 9 * Cin;
        hipLaunchKernelGGL(column_sum, dim3((unsigned)((n + 63) / 64)), dim3(1024), 0, s, (const float*)work, dw, n,
                           pl.S);
        return finish(hipSuccess);
    }
    a.M = Cout;
    a.N = 9 * Cin;
    a.K = B * H * W;
    const bool vec = vec_ok(Cin, Cout);
    const int S = wgrad_splits(a.K, a.M, a.N, vec);
    a.kchunk = wgrad_chunk(a.K, S);
    const dim3 g((unsigned)((a.M + BM - 1) / BM), (unsigned)((a.N + BN - 1) / BN), (unsigned)S);
    if (vec) launch_gemm2<MODE_WGRAD>(a, S, s);
    else hipLaunchKernelGGL((conv_gemm<MODE_WGRAD, false>), g, dim3(256), 0, s, a);
    const int64_t n = (int64_t)a.M * a.N;
    hipLaunchKernelGGL(column_sum, dim3((unsigned)((n + 63) / 64)), dim3(1024), 0, s, (const float*)work, dw, n, S);
    return finish(hipSuccess);
}

int mpd_nn_bias_grad(const float* dz, float* db, float* work, int P, int C, void* stream) {
    if (!dz || !db || !work || P <= 0 || C <= 0) return bad("bias_grad: bad argument");
    const int S = bias_slices(P, C);
    const int rows = (P + S - 1) / S;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(bias_grad_part, dim3((unsigned)((C + bias_cw(C) - 1) / bias_cw(C)), (unsigned)S), dim3(256), 0,
                       s, dz, work, P, C, rows);
    hipLaunchKernelGGL(column_sum, dim3((unsigned)((C + 63) / 64)), dim3(1024), 0, s, (const float*)work, db, (int64_t)C, S);
    return finish(hipSuccess);
}

int mpd_nn_silu_bwd_bias(const float* z, const float* da, uint16_t* dz_bf16, float* db, float* work, int P, int C,
                         void* stream) {
    if (!da || !dz_bf16 || !db || !work || P <= 0 || C <= 0) return bad("silu_bwd_bias: bad argument");
    const int S = bias_slices(P, C);
    const int rows = (P + S - 1) / S;
    hipStream_t s = (hipStream_t)stream;
    const dim3 g((unsigned)((C + bias_cw(C) - 1) / bias_cw(C)), (unsigned)S);
    if (z) hipLaunchKernelGGL(silu_bwd_bias_part<true>, g, dim3(256), 0, s, z, da, dz_bf16, work, P, C, rows);
    else hipLaunchKernelGGL(silu_bwd_bias_part<false>, g, dim3(256), 0, s, z, da, dz_bf16, work, P, C, rows);
    hipLaunchKernelGGL(column_sum, dim3((unsigned)((C + 63) / 64)), dim3(1024), 0, s, (const float*)work, db, (int64_t)C, S);
    return finish(hipSuccess);
}

long long mpd_nn_bias_work_floats(int P, int C) {
    if (P <= 0 || C <= 0) return -1;
    return (long long)bias_slices(P, C) * C;
}

int mpd_nn_silu_fwd(const float* z, uint16_t* a, long long n, void* stream) {
    if (!z || !a || n < 0) return bad("silu_fwd: bad argument");
    hipLaunchKernelGGL(silu_fwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, z, a, (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_silu_bwd(const float* z, const float* da, uint16_t* dz_bf16, float* dz_f32, long long n, void* stream) {
    if (!z || !da || !dz_bf16 || n < 0) return bad("silu_bwd: bad argument");
    hipLaunchKernelGGL(silu_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, z, da, dz_bf16, dz_f32,
                       (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_to_bf16(const float* x, uint16_t* y, long long n, void* stream) {
    if (!x || !y || n < 0) return bad("to_bf16: bad argument");
    hipLaunchKernelGGL(to_bf16, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_avgpool2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream) {
    if (!x || !y || B <= 0 || H < 2 || W < 2 || C <= 0 || (H & 1) || (W & 1)) return bad("avgpool2_fwd: bad argument");
    const int64_t n = (int64_t)B * (H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(avgpool2_fwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B, H, W, C);
    return finish(hipSuccess);
}

int mpd_nn_avgpool2_bwd(const float* dy, float* dx, int B, int H, int W, int C, void* stream) {
    if (!dy || !dx || B <= 0 || H < 2 || W < 2 || C <= 0 || (H & 1) || (W & 1)) return bad("avgpool2_bwd: bad argument");
    const int64_t n = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(avgpool2_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dy, dx, B, H, W, C);
    return finish(hipSuccess);
}

int mpd_nn_upsample2_fwd(const uint16_t* x, uint16_t* y, int B, int H, int W, int C, void* stream) {
    if (!x || !y || B <= 0 || H <= 0 || W <= 0 || C <= 0) return bad("upsample2_fwd: bad argument");
    const int64_t n = (int64_t)B * 4 * H * W * C;
    hipLaunchKernelGGL(upsample2_fwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, B, H, W, C);
    return finish(hipSuccess);
}

int mpd_nn_upsample2_bwd(const float* dy, float* dx, int B, int H, int W, int C, void* stream) {
    if (!dy || !dx || B <= 0 || H <= 0 || W <= 0 || C <= 0) return bad("upsample2_bwd: bad argument");
    const int64_t n = (int64_t)B * H * W * C;
    hipLaunchKernelGGL(upsample2_bwd, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, dy, dx, B, H, W, C);
    return finish(hipSuccess);
}

int mpd_nn_concat_fwd(const uint16_t* a, int Ca, const uint16_t* b, int Cb, uint16_t* y, long long P, void* stream) {
    if (!a || !b || !y || Ca <= 0 || Cb <= 0 || P < 0) return bad("concat_fwd: bad argument");
    hipLaunchKernelGGL(concat_fwd, dim3(blocks_for(P * (Ca + Cb))), dim3(256), 0, (hipStream_t)stream, a, Ca, b, Cb,
                       y, (int64_t)P);
    return finish(hipSuccess);
}

int mpd_nn_concat_bwd(const float* dy, float* da, int Ca, int acc_a, float* db, int Cb, int acc_b, long long P,
                      void* stream) {
    if (!dy || !da || !db || Ca <= 0 || Cb <= 0 || P < 0) return bad("concat_bwd: bad argument");
    hipLaunchKernelGGL(concat_bwd, dim3(blocks_for(P * (Ca + Cb))), dim3(256), 0, (hipStream_t)stream, dy, da, Ca,
                       acc_a, db, Cb, acc_b, (int64_t)P);
    return finish(hipSuccess);
}

int mpd_nn_upcat_fwd(const uint16_t* x, int Cu, const uint16_t* skip, int Cs, uint16_t* y, int B, int H, int W,
                     void* stream) {
    const long long runs = 4LL * B * H * W * ((Cu + Cs) / 8);
    if (!x || !skip || !y || B <= 0 || H <= 0 || W <= 0 || Cu <= 0 || Cs <= 0 || Cu % 8 || Cs % 8 ||
        runs >= (1LL << 31))
        return bad("upcat_fwd: bad argument (channel counts must be positive multiples of 8)");
    hipLaunchKernelGGL(upcat_fwd, dim3(blocks_for(runs)), dim3(256), 0, (hipStream_t)stream, x, Cu, skip, Cs, y, B, H,
                       W);
    return finish(hipSuccess);
}

int mpd_nn_upcat_bwd(const float* dy, float* dx, int Cu, float* dskip, int Cs, int B, int H, int W, void* stream) {
    const long long runs = 1LL * B * H * W * ((Cu + Cs) / 4);
    if (!dy || !dx || !dskip || B <= 0 || H <= 0 || W <= 0 || Cu <= 0 || Cs <= 0 || Cu % 4 || Cs % 4 ||
        4LL * B * H * W * (Cu + Cs) >= (1LL << 31))
        return bad("upcat_bwd: bad argument (channel counts must be positive multiples of 4)");
    hipLaunchKernelGGL(upcat_bwd, dim3(blocks_for(runs)), dim3(256), 0, (hipStream_t)stream, dy, dx, Cu, dskip, Cs, B,
                       H, W);
    return finish(hipSuccess);
}

int mpd_nn_pool_bwd_add(const float* dy, const float* skip, float* dx, int B, int H, int W, int C, void* stream) {
    const long long runs = 1LL * B * H * W * (C / 4);
    if (!dy || !skip || !dx || B <= 0 || H <= 0 || W <= 0 || H % 2 || W % 2 || C <= 0 || C % 4 ||
        1LL * B * H * W * C >= (1LL << 31))
        return bad("pool_bwd_add: bad argument (H, W even, C a positive multiple of 4)");
    hipLaunchKernelGGL(pool_bwd_add, dim3(blocks_for(runs)), dim3(256), 0, (hipStream_t)stream, dy, skip, dx, B, H, W,
                       C);
    return finish(hipSuccess);
}

int mpd_nn_pool_silu_bwd_bias(const float* dy, const float* skip, const float* z, uint16_t* dz_bf16, float* db,
                              float* work, int B, int H, int W, int C, void* stream) {
    if (!dy || !skip || !z || !dz_bf16 || !db || !work || B <= 0 || H <= 0 || W <= 0 || H % 2 || W % 2 || C <= 0 ||
        1LL * B * H * W * C >= (1LL << 31))
        return bad("pool_silu_bwd_bias: bad argument (H, W even, B H W C below 2^31)");
    const int P = B * H * W;
    const int S = bias_slices(P, C);
    const int rows = (P + S - 1) / S;
    hipStream_t s = (hipStream_t)stream;
    const DivMagic wd = div_magic(W);
    hipLaunchKernelGGL(pool_silu_bwd_bias_part, dim3((unsigned)((C + bias_cw(C) - 1) / bias_cw(C)), (unsigned)S),
                       dim3(256), 0, s, dy, skip, z, dz_bf16, work, P, C, W, wd.m, wd.sh, rows);
    hipLaunchKernelGGL(column_sum, dim3((unsigned)((C + 63) / 64)), dim3(1024), 0, s, (const float*)work, db, (int64_t)C, S);
    return finish(hipSuccess);
}

int mpd_nn_add(float* y, const float* x, long long n, void* stream) {
    if (!y || !x || n < 0) return bad("add: bad argument");
    hipLaunchKernelGGL(add_inplace, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, y, x, (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_mse_blocks(long long n) { return (int)std::min<long long>(1024, std::max<long long>(1, (n + 255) / 256)); }

int mpd_nn_mse(const float* y, const float* t, float* dy, float* part, long long n, void* stream) {
    if (!y || !t || !part || n <= 0) return bad("mse: bad argument");
    hipLaunchKernelGGL(mse, dim3((unsigned)mpd_nn_mse_blocks(n)), dim3(256), 0, (hipStream_t)stream, y, t, dy, part,
                       (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_adam(float* w, uint16_t* w_bf16, const float* g, float* m, float* v, long long n, float lr, float b1,
                float b2, float eps, int step, float gscale, void* stream) {
    if (!w || !g || !m || !v || n < 0 || step < 1) return bad("adam: bad argument");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, w, w_bf16, g, m, v, (int64_t)n,
                       lr, b1, b2, eps, c1, c2, gscale);
    return finish(hipSuccess);
}

int mpd_nn_adam_multi(int count, float* const* w, uint16_t* const* w_bf16, const float* const* g, float* const* m,
                      float* const* v, const long long* n, float lr, float b1, float b2, float eps, int step,
                      float gscale, void* stream) {
    if (count < 0 || (count > 0 && (!w || !g || !m || !v || !n)) || step < 1) return bad("adam_multi: bad argument");
    for (int k = 0; k < count; ++k)
        if (!w[k] || !g[k] || !m[k] || !v[k] || n[k] < 0) return bad("adam_multi: bad argument");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    for (int k0 = 0; k0 < count; k0 += kAdamMax) {
        AdamTable t{};
        t.count = std::min(kAdamMax, count - k0);
        int64_t blocks = 0;
        for (int j = 0; j < t.count; ++j) {
            const int k = k0 + j;
            t.s[j] = AdamSeg{w[k], w_bf16 ? w_bf16[k] : nullptr, g[k], m[k], v[k], (int64_t)n[k], blocks};
            blocks += (n[k] + 1023) / 1024;
        }
        if (blocks == 0) continue;
        if (blocks > 0x7fffffffLL) return bad("adam_multi: too many elements");
        hipLaunchKernelGGL(adam_multi, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, t, lr, b1, b2, eps,
                           c1, c2, gscale);
    }
    return finish(hipSuccess);
}

int mpd_nn_sumsq_blocks(long long n) { return (int)std::min<long long>(64, std::max<long long>(1, (n + 255) / 256)); }

int mpd_nn_sumsq(const float* x, float* part, long long n, void* stream) {
    if (!x || !part || n <= 0) return bad("sumsq: bad argument");
    hipLaunchKernelGGL(sumsq, dim3((unsigned)mpd_nn_sumsq_blocks(n)), dim3(256), 0, (hipStream_t)stream, x, part,
                       (int64_t)n);
    return finish(hipSuccess);
}

int mpd_nn_gather_rows(const float* src, const int* idx, float* dst, int nrows, long long row_len, void* stream) {
    if (!src || !idx || !dst || nrows < 0 || row_len <= 0) return bad("gather_rows: bad argument");
    hipLaunchKernelGGL(gather_rows, dim3(blocks_for((int64_t)nrows * row_len)), dim3(256), 0, (hipStream_t)stream, src,
                       idx, dst, nrows, (int64_t)row_len);
    return finish(hipSuccess);
}

}  // extern "C"
