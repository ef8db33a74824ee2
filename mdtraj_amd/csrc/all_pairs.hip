// All-pairs RMSD matrix (gfx950).
//
// Layout in HBM (workspace): every frame's selection is centred in FP64 and
// rounded once to float32, stored axis-major per frame:
//   planes[f][axis][k]  f < F, axis < 3, k < K_pad (zero padded)
// with its trace G_f in FP64.  The correlation matrix of a pair is
// M_ij[r][c] = sum_k P_r[i,k] P_c[j,k]: nine plane products, i.e. the
// 3F x 3F Gram matrix of the stacked planes.  Only tiles on or above the
// diagonal are computed; the QCP solve (FP64) runs in the tile epilogue, one
// lane per pair, and the result goes to the condensed upper triangle.
//
// Split-precision modes store, per frame f, 32-atom chunk c, axis a and part
// p (0 hi, 1 lo), chunk-major so that one (frame, chunk) is 384 contiguous bytes,
//   split[((f * (K_pad / 32) + c) * 6 + 2 a + p) * 32 + k % 32]
// as x 2^s_f in FP16 or BF16 (a per-frame power of two), scale[f] = 2^-s_f,
// and traces_split[f] = sum (hi + lo)^2 2^-2s_f in FP64 — the trace of exactly
// the structure the matrix products see.
//   three-product form: hi = round_half(x 2^s), lo = round_half(x 2^s - hi),
//     M ~ hi.hi + hi.lo + lo.hi in one FP32 accumulator per entry;
//   exact-hi form: hi = rint(x 2^s) is an integer with |hi| <= 2^b, where
//     2 b + log2(K_pad) <= 24, so every partial sum of hi.hi is an integer
//     below 2^24 and its FP32 accumulation is exact; lo = round_half(x 2^s - hi)
//     and hi.lo + lo.hi + lo.lo go to a second accumulator whose magnitude is
//     2^-b of the first, so its FP32 rounding stays far below the FP64 bar.
#include "common.h"

namespace hr {
namespace {

constexpr int kPrepBlock = 256;
constexpr int kChunk = 32;  // atoms per chunk of the split layout (K of one MFMA)

enum class Split { ThreeProduct, ExactHi };

template <typename Half>
__device__ __forceinline__ float round_half(float x) {
    return (float)(Half)x;
}

// Gather + centre + round + trace, one wave per frame; split planes too.
// Frames of up to kPrepRegs * 64 padded atoms are loaded into registers in
// one batch (every load issued before any is used) and the three passes
// read them there; longer frames are read from memory in each pass.  Either
// way each lane sums its atoms j = lane, lane + 64, ... in that order, so the
// results do not depend on the path.
constexpr int kPrepRegs = 8;

template <typename Half, Split kSplit>
__global__ __launch_bounds__(kPrepBlock) void pairs_prep_kernel(
    const float* __restrict__ xyz, long long n_frames, int n_atoms_total,
    const int* __restrict__ idx, int n_sel, int k_pad, int hi_bits,
    float* __restrict__ planes, double* __restrict__ traces, Half* __restrict__ split,
    float* __restrict__ scale, double* __restrict__ traces_split) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (kPrepBlock / kWave) + threadIdx.x / kWave;
    const long long n_waves = (long long)gridDim.x * (kPrepBlock / kWave);
    const bool in_regs = k_pad <= kPrepRegs * kWave;  // uniform over the grid
    for (long long f = wave; f < n_frames; f += n_waves) {
        const float* fr = xyz + f * 3LL * n_atoms_total;
        float rx[kPrepRegs], ry[kPrepRegs], rz[kPrepRegs];
        if (in_regs) {
#pragma unroll
            for (int u = 0; u < kPrepRegs; ++u) {
                const int jc = min(lane + u * kWave, n_sel - 1);
                const long long a = idx ? idx[jc] : jc;
                rx[u] = fr[3 * a];
                ry[u] = fr[3 * a + 1];
                rz[u] = fr[3 * a + 2];
            }
        }
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (in_regs) {
#pragma unroll
            for (int u = 0; u < kPrepRegs; ++u) {
                if (lane + u * kWave < n_sel) {
                    sx += (double)rx[u];
                    sy += (double)ry[u];
                    sz += (double)rz[u];
                }
            }
        } else {
            for (int j = lane; j < n_sel; j += kWave) {
                const long long a = idx ? idx[j] : j;
                sx += (double)fr[3 * a];
                sy += (double)fr[3 * a + 1];
                sz += (double)fr[3 * a + 2];
            }
        }
        sx = wave_sum(sx) / n_sel;
        sy = wave_sum(sy) / n_sel;
        sz = wave_sum(sz) / n_sel;
        // centred, rounded coordinates of selected atom j (0 past the
        // selection), from memory or from register slot u
        auto centred_mem = [&](int j, float& x, float& y, float& z) {
            x = y = z = 0.0f;
            if (j < n_sel) {
                const long long a = idx ? idx[j] : j;
                x = (float)((double)fr[3 * a] - sx);
                y = (float)((double)fr[3 * a + 1] - sy);
                z = (float)((double)fr[3 * a + 2] - sz);
            }
        };
        auto centred_reg = [&](int j, int u, float& x, float& y, float& z) {
            x = y = z = 0.0f;
            if (j < n_sel) {
                x = (float)((double)rx[u] - sx);
                y = (float)((double)ry[u] - sy);
                z = (float)((double)rz[u] - sz);
            }
        };
        // runs body(j, x, y, z) over the padded atoms j = lane, lane + 64, ...
        auto each_atom = [&](auto&& body) {
            if (in_regs) {
#pragma unroll
                for (int u = 0; u < kPrepRegs; ++u) {
                    const int j = lane + u * kWave;
                    if (j < k_pad) {
                        float x, y, z;
                        centred_reg(j, u, x, y, z);
                        body(j, x, y, z);
                    }
                }
            } else {
                for (int j = lane; j < k_pad; j += kWave) {
                    float x, y, z;
                    centred_mem(j, x, y, z);
                    body(j, x, y, z);
                }
            }
        };
        float* px = planes ? planes + f * 3LL * k_pad : nullptr;
        double g = 0.0;
        float m = 0.0f;
        each_atom([&](int j, float x, float y, float z) {
            if (planes) {
                px[j] = x;
                px[k_pad + j] = y;
                px[2 * k_pad + j] = z;
            }
            const double dx = x, dy = y, dz = z;
            g = fma(dx, dx, fma(dy, dy, fma(dz, dz, g)));
            m = fmaxf(m, fmaxf(fabsf(x), fmaxf(fabsf(y), fabsf(z))));
        });
        g = wave_sum(g);
        if (lane == 0 && traces) traces[f] = g;
        if (split == nullptr) continue;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, kWave));
        int e = 0;
        if (m > 0.0f) frexpf(m, &e);  // m in [2^(e-1), 2^e)
        // three-product: max |x| 2^sh in [2^13, 2^14); exact-hi: below 2^hi_bits
        const int sh = (kSplit == Split::ExactHi ? hi_bits : 14) - e;
        const float up = ldexpf(1.0f, sh);
        const double inv = ldexp(1.0, -sh);
        double gs = 0.0;
        Half* row = split + f * 6LL * k_pad;  // this frame's k_pad / 32 chunks
        each_atom([&](int j, float x, float y, float z) {
            const float c[3] = {x, y, z};
            Half* chunk = row + (j / kChunk) * (6 * kChunk) + (j % kChunk);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float xs = c[a] * up;  // exact: power-of-two scaling
                const float hf = kSplit == Split::ExactHi ? rintf(xs) : round_half<Half>(xs);
                const float lf = round_half<Half>(xs - hf);  // xs - hf is exact in float
                chunk[(2 * a) * kChunk] = (Half)hf;
                chunk[(2 * a + 1) * kChunk] = (Half)lf;
                const double v = ((double)hf + (double)lf) * inv;
                gs = fma(v, v, gs);
            }
        });
        gs = wave_sum(gs);
        if (lane == 0) {
            scale[f] = ldexpf(1.0f, -sh);
            traces_split[f] = gs;
        }
    }
}

__device__ __forceinline__ long long condensed_index(long long i, long long j, long long n) {
    // i < j; row-major upper triangle without the diagonal
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Tile (ti, tj), ti <= tj, from a linear index over the upper triangle of a
// T x T tile grid, row by row.
__device__ __forceinline__ void triangle_tile(long long lin, long long T, long long& ti,
                                              long long& tj) {
    // row r starts at r*T - r*(r-1)/2
    double t = (double)T + 0.5;
    long long r = (long long)floor(t - sqrt(t * t - 2.0 * (double)lin));
    if (r < 0) r = 0;
    if (r > T - 1) r = T - 1;
    while (r > 0 && r * T - r * (r - 1) / 2 > lin) --r;
    while (r + 1 < T && (r + 1) * T - (r + 1) * r / 2 <= lin) ++r;
    ti = r;
    tj = r + (lin - (r * T - r * (r - 1) / 2));
}

// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs round-robin; remap so that each XCD walks a contiguous run of tiles
// (neighbouring tiles share a row block, which then stays in that XCD's L2).
__device__ __forceinline__ long long xcd_remap(long long b, long long nb) {
    constexpr int kXcd = 8;
    const long long per = (nb + kXcd - 1) / kXcd;
    const long long x = b % kXcd, q = b / kXcd;
    const long long full = nb % kXcd == 0 ? kXcd : nb % kXcd;  // XCDs with `per` blocks
    long long r;
    if (x < full) r = x * per + q;
    else r = full * per + (x - full) * (per - 1) + q;
    return r;
}

// ---------------------------------------------------------------- FP32 kernel
constexpr int kF32Tile = 64;   // frames per tile side
constexpr int kF32Block = 256; // 16 x 16 threads, 4 x 4 pairs each
constexpr int kF32K = 16;      // atoms per LDS stage

__global__ __launch_bounds__(kF32Block) void pairs_fp32_kernel(
    const float* __restrict__ planes, const double* __restrict__ traces, long long n_frames,
    int n_sel, int k_pad, long long n_tiles, float* __restrict__ out) {
    __shared__ float sA[3][kF32K][kF32Tile + 1];
    __shared__ float sB[3][kF32K][kF32Tile + 1];
    const long long T = (n_frames + kF32Tile - 1) / kF32Tile;
    const long long lin = xcd_remap(blockIdx.x, n_tiles);
    long long ti, tj;
    triangle_tile(lin, T, ti, tj);
    const long long i0 = ti * kF32Tile, j0 = tj * kF32Tile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;

    float acc[4][4][9];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int e = 0; e < 9; ++e) acc[u][v][e] = 0.0f;

    for (int k0 = 0; k0 < k_pad; k0 += kF32K) {
        // stage 64 frames x 16 atoms x 3 axes for both tiles
        for (int t = threadIdx.x; t < 3 * kF32K * kF32Tile; t += kF32Block) {
            const int k = t % kF32K;
            const int fr = (t / kF32K) % kF32Tile;
            const int ax = t / (kF32K * kF32Tile);
            const long long fi = i0 + fr, fj = j0 + fr;
            sA[ax][k][fr] = fi < n_frames ? planes[(fi * 3 + ax) * k_pad + k0 + k] : 0.0f;
            sB[ax][k][fr] = fj < n_frames ? planes[(fj * 3 + ax) * k_pad + k0 + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kF32K; ++k) {
            float a[4][3], b[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    a[u][ax] = sA[ax][k][ty + 16 * u];
                    b[u][ax] = sB[ax][k][tx + 16 * u];
                }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            acc[u][v][3 * r + c] = fmaf(a[u][r], b[v][c], acc[u][v][3 * r + c]);
        }
        __syncthreads();
    }
    const double n = (double)n_sel;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = i0 + ty + 16 * u;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long long j = j0 + tx + 16 * v;
            if (i < n_frames && j < n_frames && i < j) {
                double M[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) M[e] = (double)acc[u][v][e];
                const double msd = qcp_msd(M, traces[i], traces[j], n);
                out[condensed_index(i, j, n_frames)] = (float)sqrt(msd);
            }
        }
    }
}

// ------------------------------------------------------- split-MFMA kernel
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename Half> struct Mfma;
template <> struct Mfma<_Float16> {
    typedef f16x8 vec;
    static __device__ __forceinline__ f32x4 run(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<__bf16> {
    typedef bf16x8 vec;
    static __device__ __forceinline__ f32x4 run(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// Workgroup: 8 waves (2 per SIMD) on a 64 x 64 frame tile; wave w owns rows
// 32 (w / 4) .. +32 (two 16-row MFMA blocks) and columns 16 (w % 4) .. +16.
// Operands stream through a 3-stage LDS ring filled by direct global->LDS
// loads (global_load_lds_dwordx4): a stage is one 32-atom chunk of both tiles,
// each frame row 24 16-byte slots (6 planes x 4 quarters) padded to 25, which
// makes the 16-lane operand reads bank-conflict free.
constexpr int kMBlock = 512;
constexpr int kMTile = 64;
constexpr int kSlots = 25;                       // 24 data + 1 pad, 16 B each
constexpr int kTileBytes = kMTile * kSlots * 16; // 25,600
constexpr int kStageBytes = 2 * kTileBytes;      // A and B tiles
constexpr int kStages = 3;
constexpr size_t kMfmaLds = (size_t)kStages * kStageBytes;  // 153,600
constexpr int kLoadsPerStage = 2 * kMTile * kSlots / kWave;  // 50 wave-instructions
static_assert(kLoadsPerStage == 50, "load plan assumes 50 instructions per stage");
constexpr int kSuper = 4;  // tiles per super-tile side (L2 reuse of A and B)

// Wave w issues instructions [first, first + count) of a stage: waves 0-1
// seven, waves 2-7 six (2 x 7 + 6 x 6 = 50), all lanes active, so each wave
// knows exactly how many loads of a stage it still has in flight.
__device__ __forceinline__ int loads_first(int w) { return w < 2 ? 7 * w : 14 + 6 * (w - 2); }
__device__ __forceinline__ int loads_count(int w) { return w < 2 ? 7 : 6; }

// Operand sets of one MFMA Gram launch: rows come from set A, columns from
// set B (the same set for the all-pairs triangle).
template <typename Half>
struct GramArgs {
    const Half* split_a;
    const Half* split_b;
    const float* scale_a;
    const float* scale_b;
    const double* traces_a;
    const double* traces_b;
    long long n_a, n_b;  // frames in A and B
    int n_sel, k_pad;
    long long n_blocks;
    long long tiles_b;   // rectangle: tile columns
    float* out;
    long long out_ld;    // rectangle: row stride of out
    // rectangle, nearest-column mode: instead of writing out, every row i
    // keeps min over its columns of (float bits of the RMSD) << 32 | j
    unsigned long long* row_best;
};

constexpr int kMaxLoads = 7;  // loads per wave per stage (waves 0-1: 7, the rest: 6)

// Chunk-0 source of each of this lane's stage loads.  The split layout keeps
// a frame's chunks 6 * kChunk halves (384 bytes) apart, so the load for chunk
// c reads src[k] + c * 384 bytes: the per-tile address arithmetic runs once,
// not once per chunk.  Entries k >= the wave's load count are never issued.
template <typename Half>
__device__ __forceinline__ void plan_loads(const GramArgs<Half>& g, long long i0, long long j0,
                                           int n_chunks, int first, int lane,
                                           const char* (&src)[kMaxLoads]) {
#pragma unroll
    for (int k = 0; k < kMaxLoads; ++k) {
        const int slot = (first + k) * kWave + lane;
        const int tile = slot >= kMTile * kSlots;  // 0: A rows, 1: B columns
        const int sl = slot - tile * (kMTile * kSlots);
        const int fr = min(sl / kSlots, kMTile - 1);  // slots past 2 tiles: never issued
        int r = sl - fr * kSlots;
        if (r >= kSlots - 1) r = 0;  // pad slot: harmless duplicate load
        const long long nf = tile ? g.n_b : g.n_a;
        long long f = (tile ? j0 : i0) + fr;
        if (f >= nf) f = nf - 1;     // rows past the end: never written out
        const Half* base = tile ? g.split_b : g.split_a;
        src[k] = reinterpret_cast<const char*>(base + (f * n_chunks * 6) * kChunk + r * 8);
    }
}

// DMA of chunk c into one LDS stage: 16 bytes per lane per load, straight
// from global memory to LDS (no staging registers).  count is wave-uniform.
template <typename Half>
__device__ __forceinline__ void issue_stage(const char* const (&src)[kMaxLoads], int first,
                                            int count, int c, char* stage) {
    const long long adv = (long long)c * 6 * kChunk * (long long)sizeof(Half);
#pragma unroll
    for (int k = 0; k < kMaxLoads; ++k) {
        if (k < count)
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const void*>(src[k] + adv),
                (__attribute__((address_space(3))) void*)(stage + (first + k) * kWave * 16), 16,
                0, 0);
    }
}

__device__ __forceinline__ void wait_loads(int w, bool keep_one) {
    // s_waitcnt encoding (gfx9): vmcnt[3:0] | expcnt 6:4 | lgkmcnt 11:8 | vmcnt[5:4] 15:14
    if (!keep_one) __builtin_amdgcn_s_waitcnt(0x0F70);
    else if (w < 2) __builtin_amdgcn_s_waitcnt(0x0F70 | 7);
    else __builtin_amdgcn_s_waitcnt(0x0F70 | 6);
}

// Tile of block b: super-tiles of kSuper x kSuper tiles, upper triangle of
// super-tiles row by row, tiles row-major inside; false for empty slots.
__device__ __forceinline__ bool super_tile(long long b, long long T, long long& ti,
                                           long long& tj) {
    const long long TS = (T + kSuper - 1) / kSuper;
    const long long s = b / (kSuper * kSuper);
    const int in = (int)(b - s * (kSuper * kSuper));
    long long si, sj;
    triangle_tile(s, TS, si, sj);
    ti = si * kSuper + in / kSuper;
    tj = sj * kSuper + in % kSuper;
    return ti <= tj && tj < T;
}

// Accumulators acc[set][3 r + c][rb] (16 x 16 blocks) collect P_r[i] . P_c[j]
// over the atoms; the exact-hi form keeps hi.hi in set 0 and the cross
// products in set 1.  16x16x32 layouts: operand lane l holds row (column)
// l % 16, k = 8 (l / 16) .. +8; result register v of lane l holds row
// 4 (l / 16) + v, column l % 16.
//   kRect false: upper triangle of A x A into the condensed vector;
//   kRect true:  full A x B rectangle, out[i * out_ld + j].
template <typename Half, Split kSplit, bool kRect>
__global__ __launch_bounds__(kMBlock) void pairs_mfma_kernel(const GramArgs<Half> g) {
    typedef typename Mfma<Half>::vec vec;
    constexpr bool kExact = kSplit == Split::ExactHi;
    constexpr int kSets = kExact ? 2 : 1;
    extern __shared__ __attribute__((aligned(16))) char lds[];

    long long ti, tj;
    const long long lin = xcd_remap(blockIdx.x, g.n_blocks);
    if constexpr (kRect) {
        ti = lin / g.tiles_b;
        tj = lin - ti * g.tiles_b;
    } else {
        const long long T = (g.n_a + kMTile - 1) / kMTile;
        if (!super_tile(lin, T, ti, tj)) return;  // whole block
    }
    const long long i0 = ti * kMTile, j0 = tj * kMTile;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int wi = w >> 2, wj = w & 3;
    // triangle, diagonal tile: rows 32 wi.. vs columns 16 wj..: the wave has
    // work unless its whole block lies below the diagonal
    const bool active = kRect || !(ti == tj && 32 * wi >= 16 * wj + 16);
    const int n_chunks = g.k_pad / kChunk;

    f32x4 acc[kSets][9][2];
#pragma unroll
    for (int st = 0; st < kSets; ++st)
#pragma unroll
        for (int e = 0; e < 9; ++e)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) acc[st][e][rb] = f32x4{0.f, 0.f, 0.f, 0.f};

    // LDS byte offsets of this lane's operand rows (plane q adds 64 q)
    const int quarter = lane >> 4;
    int aoff[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
        aoff[rb] = (32 * wi + 16 * rb + (lane & 15)) * kSlots * 16 + quarter * 16;
    const int boff = kTileBytes + (16 * wj + (lane & 15)) * kSlots * 16 + quarter * 16;

    // w is wave-uniform; readfirstlane lets the compiler keep the load
    // count and the LDS offsets in scalar registers
    const int wu = __builtin_amdgcn_readfirstlane(w);
    const int first = loads_first(wu), count = loads_count(wu);
    const char* src[kMaxLoads];
    plan_loads<Half>(g, i0, j0, n_chunks, first, lane, src);
    issue_stage<Half>(src, first, count, 0, lds);
    if (n_chunks > 1) issue_stage<Half>(src, first, count, 1, lds + kStageBytes);
    for (int c = 0; c < n_chunks; ++c) {
        wait_loads(w, c + 1 < n_chunks);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this wave's LDS reads are done
        __builtin_amdgcn_s_barrier();        // every wave's chunk-c loads have landed
        if (c + 2 < n_chunks)
            issue_stage<Half>(src, first, count, c + 2,
                              lds + ((c + 2) % kStages) * kStageBytes);
        if (!active) continue;
        const char* st = lds + (c % kStages) * kStageBytes;
        vec b[6];  // planes q = 2 axis + part of this wave's 16 columns
#pragma unroll
        for (int q = 0; q < 6; ++q) b[q] = *reinterpret_cast<const vec*>(st + boff + q * 64);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            vec a[2][2];  // [rb][part] of plane r
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int p = 0; p < 2; ++p)
                    a[rb][p] = *reinterpret_cast<const vec*>(st + aoff[rb] + (2 * r + p) * 64);
            constexpr int X = kSets - 1;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[0][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][0], b[2 * cc], acc[0][3 * r + cc][rb]);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[X][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][0], b[2 * cc + 1], acc[X][3 * r + cc][rb]);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[X][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][1], b[2 * cc], acc[X][3 * r + cc][rb]);
            if constexpr (kExact) {
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
                        acc[X][3 * r + cc][rb] =
                            Mfma<Half>::run(a[rb][1], b[2 * cc + 1], acc[X][3 * r + cc][rb]);
            }
        }
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __syncthreads();

    // Epilogue, two rounds (rb = 0, 1): every wave parks its accumulators of
    // one 16-row block in LDS ([wave][set][entry][v][lane] floats, at most
    // 8 x 2 x 9 x 4 x 64 x 4 B = 147,456 B), then all 512 threads solve the
    // 2,048 pairs of the round, four each, two at a time with their QCP
    // solves interleaved (qcp_msd2).
    float* ep = reinterpret_cast<float*>(lds);
    const double n = (double)g.n_sel;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if (active) {
#pragma unroll
            for (int st = 0; st < kSets; ++st)
#pragma unroll
                for (int e = 0; e < 9; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        ep[(((w * kSets + st) * 9 + e) * 4 + v) * 64 + lane] = acc[st][e][rb][v];
        }
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < 4; k += 2) {
            // two pairs per step, their QCP solves interleaved
            double M2[2][9], ga[2], gb[2];
            long long o2[2];
            bool ok[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int p = threadIdx.x + kMBlock * (k + h);  // 0..2047
                const int pw = p >> 8, v = (p >> 6) & 3, ln = p & 63;
                const int pwi = pw >> 2, pwj = pw & 3;
                const bool live = kRect || !(ti == tj && 32 * pwi >= 16 * pwj + 16);
                const long long i = i0 + 32 * pwi + 16 * rb + 4 * (ln >> 4) + v;
                const long long j = j0 + 16 * pwj + (ln & 15);
                ok[h] = live && (kRect ? (i < g.n_a && j < g.n_b) : (i < j && j < g.n_a));
                const long long ic = ok[h] ? i : 0, jc = ok[h] ? j : 0;
                const double sij = (double)(g.scale_a[ic] * g.scale_b[jc]);
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    double val = (double)ep[(((pw * kSets + 0) * 9 + e) * 4 + v) * 64 + ln];
                    if constexpr (kExact)
                        val += (double)ep[(((pw * kSets + 1) * 9 + e) * 4 + v) * 64 + ln];
                    M2[h][e] = val * sij;
                }
                ga[h] = g.traces_a[ic];
                gb[h] = g.traces_b[jc];
                o2[h] = ok[h] ? (kRect ? i * g.out_ld + j : condensed_index(i, j, g.n_a)) : 0;
                if (!ok[h]) {
                    // a pair not written out (its LDS slot may be stale): a
                    // small well-conditioned stand-in (M = I, lambda = 3)
#pragma unroll
                    for (int e = 0; e < 9; ++e) M2[h][e] = (e % 4 == 0) ? 1.0 : 0.0;
                    ga[h] = 3.5;
                    gb[h] = 3.5;
                }
            }
            const bool any = ok[0] || ok[1];
            if (!g.row_best && !any) continue;
            double msd[2] = {0.0, 0.0};
            if (any) qcp_msd2(M2[0], ga[0], gb[0], M2[1], ga[1], gb[1], n, msd[0], msd[1]);
            if (!g.row_best) {
                // one float rounding, then the float root
                if (ok[0]) g.out[o2[0]] = sqrtf((float)msd[0]);
                if (ok[1]) g.out[o2[1]] = sqrtf((float)msd[1]);
                continue;
            }
            // nearest column of each row: the 16 lanes of a lane group share
            // the row (i) and hold 16 consecutive columns (j); an RMSD is >= 0,
            // so its float bits order like its value, and the column in the
            // low word makes ties go to the lowest column
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int p = threadIdx.x + kMBlock * (k + h);
                const int pw = p >> 8, v = (p >> 6) & 3, ln = p & 63;
                const long long i = i0 + 32 * (pw >> 2) + 16 * rb + 4 * (ln >> 4) + v;
                const long long j = j0 + 16 * (pw & 3) + (ln & 15);
                unsigned long long key =
                    ok[h] ? ((unsigned long long)__float_as_uint(sqrtf((float)msd[h])) << 32) |
                                (unsigned long long)(unsigned)j
                          : ~0ull;
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) {
                    const unsigned hi = __shfl_xor((unsigned)(key >> 32), m, kWave);
                    const unsigned lo = __shfl_xor((unsigned)key, m, kWave);
                    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
                    key = o < key ? o : key;
                }
                if ((ln & 15) == 0 && key != ~0ull) atomicMin(g.row_best + i, key);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------- split-MFMA kernel, half tiles
// The same Gram and epilogue on 64 x 32 pair tiles: 4 waves (one per SIMD)
// of 32 x 16 pairs each, and a 2-stage operand ring of 77,824 bytes, so that
// two workgroups fit on a CU (2 waves per SIMD, at most 256 registers each).
// The workgroups of a CU run independently, so one's FP64 QCP epilogue
// overlaps the other's MFMA loop, where the 64 x 64 kernel (one workgroup
// per CU) leaves the matrix cores idle during its epilogue.  Both halves of a
// 64 x 64 tile come from consecutive blocks, so after the XCD remap they
// share their 64 A rows in one L2.
constexpr int kHBlock = 256;
constexpr int kHRowsA = 64, kHRowsB = 32;
constexpr int kHStages = 2;

// Operand ring of the half-tile kernel for frame rows of kS 16-byte slots
// (24 data + padding): data slots, load instructions per stage (the last
// one's lanes past the data land in the stage's slack), bytes.
template <int kS>
struct HalfRing {
    static constexpr int kSlotsData = (kHRowsA + kHRowsB) * kS;
    static constexpr int kLoads = (kSlotsData + kWave - 1) / kWave;
    static constexpr int kStageBytes = kLoads * kWave * 16;
    static constexpr size_t kLds = (size_t)kHStages * kStageBytes;
    static_assert(kLds <= 81920, "two workgroups per CU must fit 160 KiB of LDS");
    static_assert(4 * 2 * 9 * 4 * kWave * 4 <= (int)kLds, "epilogue park must fit the ring");
    // waves 0 .. kLoads % 4 - 1 issue one instruction more than the others
    static __device__ __forceinline__ int first(int w) {
        const int base = kLoads / 4, extra = kLoads % 4;
        return w * base + (w < extra ? w : extra);
    }
    static __device__ __forceinline__ int count(int w) {
        return kLoads / 4 + (w < kLoads % 4 ? 1 : 0);
    }
};
constexpr int kHMaxLoads = 10;
static_assert(HalfRing<25>::kLoads == 38 && HalfRing<26>::kLoads == 39, "load plans");
static_assert(HalfRing<26>::kLds == 79872, "dynamic LDS of the 26-slot ring");

template <typename Half, int kS>
__device__ __forceinline__ void plan_loads_half(const GramArgs<Half>& g, long long i0,
                                                long long j0, int n_chunks, int first, int lane,
                                                const char* (&src)[kHMaxLoads]) {
#pragma unroll
    for (int k = 0; k < kHMaxLoads; ++k) {
        int slot = (first + k) * kWave + lane;
        if (slot >= HalfRing<kS>::kSlotsData) slot = 0;  // slack lanes: a valid duplicate source
        const int tile = slot >= kHRowsA * kS;           // 0: A rows, 1: B columns
        const int sl = slot - tile * (kHRowsA * kS);
        const int fr = sl / kS;
        int r = sl - fr * kS;
        if (r >= 24) r = 0;                              // pad slots: harmless duplicate loads
        const long long nf = tile ? g.n_b : g.n_a;
        long long f = (tile ? j0 : i0) + fr;
        if (f >= nf) f = nf - 1;                    // rows past the end: never written out
        const Half* base = tile ? g.split_b : g.split_a;
        src[k] = reinterpret_cast<const char*>(base + (f * n_chunks * 6) * kChunk + r * 8);
    }
}

template <typename Half>
__device__ __forceinline__ void issue_stage_half(const char* const (&src)[kHMaxLoads], int first,
                                                 int count, int c, char* stage) {
    const long long adv = (long long)c * 6 * kChunk * (long long)sizeof(Half);
#pragma unroll
    for (int k = 0; k < kHMaxLoads; ++k) {
        if (k < count)
            __builtin_amdgcn_global_load_lds(
                reinterpret_cast<const void*>(src[k] + adv),
                (__attribute__((address_space(3))) void*)(stage + (first + k) * kWave * 16), 16,
                0, 0);
    }
}

// Tile of block b: the 64 x 64 super-tile order of the full kernel, each
// 64 x 64 tile split into two column halves taken by consecutive blocks.
//   kRect false: upper triangle of A x A into the condensed vector;
//   kRect true:  the A x B rectangle, tiles_b 32-column tiles per row block.
//   kS: 16-byte slots per frame row in LDS (25 or 26; which one reads the
//       operands without bank conflicts depends on the LDS's bank count).
//   kRegEpi: every lane solves the QCP of the 8 pairs its accumulators hold,
//       straight from registers; otherwise the accumulators are parked in LDS
//       and all 256 threads share the solves.
template <typename Half, Split kSplit, bool kRect, int kS, bool kRegEpi>
__global__ __launch_bounds__(kHBlock, 2) void pairs_mfma_half_kernel(const GramArgs<Half> g) {
    typedef typename Mfma<Half>::vec vec;
    constexpr bool kExact = kSplit == Split::ExactHi;
    constexpr int kSets = kExact ? 2 : 1;
    extern __shared__ __attribute__((aligned(16))) char lds[];

    long long ti, tj;
    const long long lin = xcd_remap(blockIdx.x, g.n_blocks);
    long long i0, j0;
    if constexpr (kRect) {
        ti = lin / g.tiles_b;
        tj = lin - ti * g.tiles_b;
        i0 = ti * kHRowsA;
        j0 = tj * kHRowsB;
    } else {
        const long long T = (g.n_a + kMTile - 1) / kMTile;
        if (!super_tile(lin >> 1, T, ti, tj)) return;  // whole block
        i0 = ti * kMTile;
        j0 = tj * kMTile + kHRowsB * (lin & 1);
        if (j0 >= g.n_a) return;                       // whole block: a half past the end
    }
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int wi = w >> 1, wj = w & 1;
    // a wave has work unless its 32 x 16 block lies on or below the diagonal
    const bool active = kRect || j0 + 16 * wj + 15 > i0 + 32 * wi;
    const int n_chunks = g.k_pad / kChunk;

    f32x4 acc[kSets][9][2];
#pragma unroll
    for (int st = 0; st < kSets; ++st)
#pragma unroll
        for (int e = 0; e < 9; ++e)
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) acc[st][e][rb] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int quarter = lane >> 4;
    constexpr int kStageBytes = HalfRing<kS>::kStageBytes;
    int aoff[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
        aoff[rb] = (32 * wi + 16 * rb + (lane & 15)) * kS * 16 + quarter * 16;
    const int boff = kHRowsA * kS * 16 + (16 * wj + (lane & 15)) * kS * 16 + quarter * 16;

    const int wu = __builtin_amdgcn_readfirstlane(w);
    const int first = HalfRing<kS>::first(wu), count = HalfRing<kS>::count(wu);
    const char* src[kHMaxLoads];
    plan_loads_half<Half, kS>(g, i0, j0, n_chunks, first, lane, src);
    issue_stage_half<Half>(src, first, count, 0, lds);
    // The two workgroups of a CU drift out of phase, so one wave of a SIMD can
    // be in the Gram loop while the other solves its FP64 epilogue: the loop
    // runs at raised issue priority so that its MFMAs are issued first and
    // the epilogue fills the gaps (measured 0.06-0.1 ms per 20,000-frame
    // matrix; the results are the same bits).
    __builtin_amdgcn_s_setprio(1);
    for (int c = 0; c < n_chunks; ++c) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's chunk-c loads landed
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): its chunk c-1 LDS reads are done
        __builtin_amdgcn_s_barrier();        // every wave: chunk c in LDS, stage c+1 free
        if (c + 1 < n_chunks)
            issue_stage_half<Half>(src, first, count, c + 1, lds + ((c + 1) & 1) * kStageBytes);
        if (!active) continue;
        const char* st = lds + (c & 1) * kStageBytes;
        vec b[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) b[q] = *reinterpret_cast<const vec*>(st + boff + q * 64);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            vec a[2][2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int p = 0; p < 2; ++p)
                    a[rb][p] = *reinterpret_cast<const vec*>(st + aoff[rb] + (2 * r + p) * 64);
            constexpr int X = kSets - 1;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[0][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][0], b[2 * cc], acc[0][3 * r + cc][rb]);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[X][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][0], b[2 * cc + 1], acc[X][3 * r + cc][rb]);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
                    acc[X][3 * r + cc][rb] =
                        Mfma<Half>::run(a[rb][1], b[2 * cc], acc[X][3 * r + cc][rb]);
            if constexpr (kExact) {
#pragma unroll
                for (int cc = 0; cc < 3; ++cc)
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
                        acc[X][3 * r + cc][rb] =
                            Mfma<Half>::run(a[rb][1], b[2 * cc + 1], acc[X][3 * r + cc][rb]);
            }
        }
    }
    __builtin_amdgcn_s_setprio(0);
    if constexpr (kRegEpi) {
        // Lane l holds, for rb and v, the nine entries of pair (row
        // 32 wi + 16 rb + 4 (l / 16) + v, column 16 wj + l % 16): its QCP runs
        // from registers, two pairs at a time (qcp_msd2), no LDS and no
        // barrier.  A wave with no pair in the triangle has nothing to do.
        if (!active && !g.row_best) return;
        const double nn = (double)g.n_sel;
        const long long j = j0 + 16 * wj + (lane & 15);
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
            for (int v0 = 0; v0 < 4; v0 += 2) {
                double M2[2][9], ga[2], gb[2];
                long long o2[2];
                bool ok[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int v = v0 + h;
                    const long long i = i0 + 32 * wi + 16 * rb + 4 * (lane >> 4) + v;
                    ok[h] = active &&
                            (kRect ? (i < g.n_a && j < g.n_b) : (i < j && j < g.n_a));
                    const long long ic = ok[h] ? i : 0, jc = ok[h] ? j : 0;
                    const double sij = (double)(g.scale_a[ic] * g.scale_b[jc]);
#pragma unroll
                    for (int e = 0; e < 9; ++e) {
                        double val = (double)acc[0][e][rb][v];
                        if constexpr (kExact) val += (double)acc[1][e][rb][v];
                        M2[h][e] = val * sij;
                    }
                    ga[h] = g.traces_a[ic];
                    gb[h] = g.traces_b[jc];
                    o2[h] = ok[h] ? (kRect ? i * g.out_ld + j : condensed_index(i, j, g.n_a)) : 0;
                    if (!ok[h]) {
                        // a pair not written out: a small well-conditioned
                        // stand-in (M = I, lambda = 3)
#pragma unroll
                        for (int e = 0; e < 9; ++e) M2[h][e] = (e % 4 == 0) ? 1.0 : 0.0;
                        ga[h] = 3.5;
                        gb[h] = 3.5;
                    }
                }
                const bool any = ok[0] || ok[1];
                double msd[2] = {0.0, 0.0};
                if (any) qcp_msd2(M2[0], ga[0], gb[0], M2[1], ga[1], gb[1], nn, msd[0], msd[1]);
                if (!g.row_best) {
                    if (ok[0]) g.out[o2[0]] = sqrtf((float)msd[0]);
                    if (ok[1]) g.out[o2[1]] = sqrtf((float)msd[1]);
                    continue;
                }
                // nearest column of each row: the 16 lanes of a lane group
                // share the row and hold 16 consecutive columns
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const long long i = i0 + 32 * wi + 16 * rb + 4 * (lane >> 4) + v0 + h;
                    unsigned long long key =
                        ok[h] ? ((unsigned long long)__float_as_uint(sqrtf((float)msd[h])) << 32) |
                                    (unsigned long long)(unsigned)j
                              : ~0ull;
#pragma unroll
                    for (int m = 1; m < 16; m <<= 1) {
                        const unsigned hi = __shfl_xor((unsigned)(key >> 32), m, kWave);
                        const unsigned lo = __shfl_xor((unsigned)key, m, kWave);
                        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
                        key = o < key ? o : key;
                    }
                    if ((lane & 15) == 0 && key != ~0ull) atomicMin(g.row_best + i, key);
                }
            }
        }
        return;
    }
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __syncthreads();

    // Epilogue, two rounds (rb = 0, 1): every wave parks one 16-row block of
    // its accumulators in LDS ([wave][set][entry][v][lane] floats, at most
    // 4 x 2 x 9 x 4 x 64 x 4 B = 73,728 B), then the 256 threads solve the
    // round's 1,024 pairs, four each, two at a time (qcp_msd2).
    float* ep = reinterpret_cast<float*>(lds);
    const double n = (double)g.n_sel;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        if (active) {
#pragma unroll
            for (int st = 0; st < kSets; ++st)
#pragma unroll
                for (int e = 0; e < 9; ++e)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        ep[(((w * kSets + st) * 9 + e) * 4 + v) * 64 + lane] = acc[st][e][rb][v];
        }
        __syncthreads();
#pragma unroll 1
        for (int k = 0; k < 4; k += 2) {
            double M2[2][9], ga[2], gb[2];
            long long o2[2];
            bool ok[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int p = threadIdx.x + kHBlock * (k + h);  // 0..1023
                const int pw = p >> 8, v = (p >> 6) & 3, ln = p & 63;
                const int pwi = pw >> 1, pwj = pw & 1;
                const bool live = kRect || j0 + 16 * pwj + 15 > i0 + 32 * pwi;
                const long long i = i0 + 32 * pwi + 16 * rb + 4 * (ln >> 4) + v;
                const long long j = j0 + 16 * pwj + (ln & 15);
                ok[h] = live && (kRect ? (i < g.n_a && j < g.n_b) : (i < j && j < g.n_a));
                const long long ic = ok[h] ? i : 0, jc = ok[h] ? j : 0;
                const double sij = (double)(g.scale_a[ic] * g.scale_b[jc]);
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    double val = (double)ep[(((pw * kSets + 0) * 9 + e) * 4 + v) * 64 + ln];
                    if constexpr (kExact)
                        val += (double)ep[(((pw * kSets + 1) * 9 + e) * 4 + v) * 64 + ln];
                    M2[h][e] = val * sij;
                }
                ga[h] = g.traces_a[ic];
                gb[h] = g.traces_b[jc];
                o2[h] = ok[h] ? (kRect ? i * g.out_ld + j : condensed_index(i, j, g.n_a)) : 0;
                if (!ok[h]) {
                    // a pair not written out (its LDS slot may be stale): a
                    // small well-conditioned stand-in (M = I, lambda = 3)
#pragma unroll
                    for (int e = 0; e < 9; ++e) M2[h][e] = (e % 4 == 0) ? 1.0 : 0.0;
                    ga[h] = 3.5;
                    gb[h] = 3.5;
                }
            }
            const bool any = ok[0] || ok[1];
            if (!g.row_best && !any) continue;
            double msd[2] = {0.0, 0.0};
            if (any) qcp_msd2(M2[0], ga[0], gb[0], M2[1], ga[1], gb[1], n, msd[0], msd[1]);
            if (!g.row_best) {
                if (ok[0]) g.out[o2[0]] = sqrtf((float)msd[0]);
                if (ok[1]) g.out[o2[1]] = sqrtf((float)msd[1]);
                continue;
            }
            // nearest column of each row, as in the 64 x 64 kernel
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int p = threadIdx.x + kHBlock * (k + h);
                const int pw = p >> 8, v = (p >> 6) & 3, ln = p & 63;
                const long long i = i0 + 32 * (pw >> 1) + 16 * rb + 4 * (ln >> 4) + v;
                const long long j = j0 + 16 * (pw & 1) + (ln & 15);
                unsigned long long key =
                    ok[h] ? ((unsigned long long)__float_as_uint(sqrtf((float)msd[h])) << 32) |
                                (unsigned long long)(unsigned)j
                          : ~0ull;
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) {
                    const unsigned hi = __shfl_xor((unsigned)(key >> 32), m, kWave);
                    const unsigned lo = __shfl_xor((unsigned)key, m, kWave);
                    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
                    key = o < key ? o : key;
                }
                if ((ln & 15) == 0 && key != ~0ull) atomicMin(g.row_best + i, key);
            }
        }
        __syncthreads();
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int k_pad_for(int n_sel) { return (n_sel + 63) / 64 * 64; }  // multiple of kChunk

// Largest b with 2 b + log2(K_pad) <= 24 (exact FP32 accumulation of hi.hi),
// capped at 7 (hi then also fits BF16's 8-bit significand exactly).
int hi_bits_for(int k_pad) {
    int lg = 0;
    while ((1 << lg) < k_pad) ++lg;
    int b = (24 - lg) / 2;
    return b > 7 ? 7 : b;
}

// Prepared split block of F frames (every piece 256-byte aligned):
//   traces F doubles | scale F floats | split F x 6 x K halfs
struct SplitBlock {
    double* traces;
    float* scale;
    void* split;
    size_t bytes;
};

SplitBlock carve_split(char* base, long long F, int K) {
    SplitBlock b{};
    size_t off = 0;
    b.traces = reinterpret_cast<double*>(base + off);
    off += align256((size_t)F * sizeof(double));
    b.scale = reinterpret_cast<float*>(base + off);
    off += align256((size_t)F * sizeof(float));
    b.split = base + off;
    off += align256((size_t)F * 6 * K * sizeof(uint16_t));
    b.bytes = off;
    return b;
}

// FP32-mode workspace: traces F doubles | planes F x 3 x K floats.
size_t fp32_ws_bytes(long long F, int K) {
    return align256((size_t)F * sizeof(double)) + align256((size_t)F * 3 * K * sizeof(float));
}

int prep_blocks_for(long long F) {
    const DeviceInfo* d = current_device();
    long long pw = (long long)(d ? d->cu_count : 256) * 16;
    if (pw > F) pw = F;
    return (int)((pw + 3) / 4);
}

template <typename Half, Split kSplit>
int launch_prep(hipStream_t s, const float* xyz, long long F, int n_atoms_total, const int* idx,
                int n_sel, int K, const SplitBlock& b) {
    const int bits = hi_bits_for(K);
    if (kSplit == Split::ExactHi && bits < 4) {
        set_error("selection of %d atoms too large for exact hi.hi accumulation", n_sel);
        return HR_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((pairs_prep_kernel<Half, kSplit>), dim3(prep_blocks_for(F)),
                       dim3(kPrepBlock), 0, s, xyz, F, n_atoms_total, idx, n_sel, K, bits,
                       static_cast<float*>(nullptr), static_cast<double*>(nullptr),
                       static_cast<Half*>(b.split), b.scale, b.traces);
    return launch_status("pairs_prep_kernel");
}

template <typename Half, Split kSplit, bool kRect>
int launch_gram(hipStream_t s, GramArgs<Half> g) {
    static bool lds_attr_set = false;
    if (!lds_attr_set) {
        HR_CHECK_HIP(hipFuncSetAttribute(
            reinterpret_cast<const void*>(pairs_mfma_kernel<Half, kSplit, kRect>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMfmaLds));
        lds_attr_set = true;
    }
    hipLaunchKernelGGL((pairs_mfma_kernel<Half, kSplit, kRect>), dim3((unsigned)g.n_blocks),
                       dim3(kMBlock), kMfmaLds, s, g);
    return launch_status("pairs_mfma_kernel");
}

template <typename Half, Split kSplit, bool kRect, int kS, bool kRegEpi>
int launch_gram_half_v(hipStream_t s, const GramArgs<Half>& g) {
    auto k = pairs_mfma_half_kernel<Half, kSplit, kRect, kS, kRegEpi>;
    static bool lds_attr_set = false;
    if (!lds_attr_set) {
        HR_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)HalfRing<kS>::kLds));
        lds_attr_set = true;
    }
    hipLaunchKernelGGL(k, dim3((unsigned)g.n_blocks), dim3(kHBlock), HalfRing<kS>::kLds, s, g);
    return launch_status("pairs_mfma_half_kernel");
}

// HR_TUNE_PAIRS_TILE: 1 the 64 x 64 kernel; 2-5 the 64 x 32 half-tile kernel
// with 25-slot rows and the LDS epilogue (2), 26-slot rows (3), the register
// epilogue (4), both (5, the default).  20,000 x 500 f16exact on MI355X, ms
// per matrix: 8.16 (1), 7.18 (2), 7.05-7.13 (3), 6.85 (4), 6.73-6.78 (5);
// every variant writes the same bits.
template <typename Half, Split kSplit, bool kRect>
int launch_gram_half(hipStream_t s, const GramArgs<Half>& g) {
    switch (tuning(HR_TUNE_PAIRS_TILE)) {
        case 2: return launch_gram_half_v<Half, kSplit, kRect, 25, false>(s, g);
        case 3: return launch_gram_half_v<Half, kSplit, kRect, 26, false>(s, g);
        case 4: return launch_gram_half_v<Half, kSplit, kRect, 25, true>(s, g);
        default: return launch_gram_half_v<Half, kSplit, kRect, 26, true>(s, g);
    }
}

bool use_half_tiles() { return tuning(HR_TUNE_PAIRS_TILE) != 1; }

// Triangle or rectangle launch over prepared sets, in the chosen tile shape.
template <typename Half, Split kSplit, bool kRect>
int launch_pairs(hipStream_t s, GramArgs<Half> g) {
    if (use_half_tiles()) {
        if constexpr (kRect) {
            g.tiles_b = (g.n_b + kHRowsB - 1) / kHRowsB;
            g.n_blocks = (g.n_a + kHRowsA - 1) / kHRowsA * g.tiles_b;
        } else {
            g.n_blocks *= 2;
        }
        if (g.n_blocks >= (1LL << 31)) {
            set_error("too many tiles (%lld)", g.n_blocks);
            return HR_ERR_UNSUPPORTED;
        }
        return launch_gram_half<Half, kSplit, kRect>(s, g);
    }
    return launch_gram<Half, kSplit, kRect>(s, g);
}

template <typename Half, Split kSplit>
int all_pairs_split(hipStream_t s, void* ws, const float* xyz, long long F, int n_atoms_total,
                    const int* idx, int n_sel, int K, float* out) {
    SplitBlock b = carve_split(static_cast<char*>(ws), F, K);
    int rc = launch_prep<Half, kSplit>(s, xyz, F, n_atoms_total, idx, n_sel, K, b);
    if (rc) return rc;
    const long long T = (F + kMTile - 1) / kMTile;
    const long long TS = (T + kSuper - 1) / kSuper;
    GramArgs<Half> g{};
    g.split_a = g.split_b = static_cast<const Half*>(b.split);
    g.scale_a = g.scale_b = b.scale;
    g.traces_a = g.traces_b = b.traces;
    g.n_a = g.n_b = F;
    g.n_sel = n_sel;
    g.k_pad = K;
    g.n_blocks = TS * (TS + 1) / 2 * kSuper * kSuper;
    g.out = out;
    if (g.n_blocks >= (1LL << 31)) {
        set_error("too many tiles (%lld)", g.n_blocks);
        return HR_ERR_UNSUPPORTED;
    }
    return launch_pairs<Half, kSplit, false>(s, g);
}

template <typename Half, Split kSplit>
int cross_split(hipStream_t s, const void* a, long long n_a, const void* b, long long n_b, int n_sel,
                int K, float* out, long long out_ld) {
    SplitBlock ba = carve_split(static_cast<char*>(const_cast<void*>(a)), n_a, K);
    SplitBlock bb = carve_split(static_cast<char*>(const_cast<void*>(b)), n_b, K);
    GramArgs<Half> g{};
    g.split_a = static_cast<const Half*>(ba.split);
    g.split_b = static_cast<const Half*>(bb.split);
    g.scale_a = ba.scale;
    g.scale_b = bb.scale;
    g.traces_a = ba.traces;
    g.traces_b = bb.traces;
    g.n_a = n_a;
    g.n_b = n_b;
    g.n_sel = n_sel;
    g.k_pad = K;
    g.tiles_b = (n_b + kMTile - 1) / kMTile;
    g.n_blocks = (n_a + kMTile - 1) / kMTile * g.tiles_b;
    g.out = out;
    g.out_ld = out_ld;
    if (g.n_blocks >= (1LL << 31)) {
        set_error("too many tiles (%lld)", g.n_blocks);
        return HR_ERR_UNSUPPORTED;
    }
    return launch_pairs<Half, kSplit, true>(s, g);
}

// Nearest column per row: keys start at all ones (a fill kernel, so the
// sequence can be captured in a graph), the rectangle kernel min-reduces
// them, and this kernel splits them into the column index and the RMSD.
__global__ __launch_bounds__(256) void fill_keys_kernel(unsigned long long* __restrict__ keys,
                                                        long long n) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) keys[i] = ~0ull;
}

__global__ __launch_bounds__(256) void unpack_keys_kernel(const unsigned long long* __restrict__ keys,
                                                          long long n, int* __restrict__ index,
                                                          float* __restrict__ value) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const unsigned long long k = keys[i];
    index[i] = (int)(unsigned)k;
    value[i] = __uint_as_float((unsigned)(k >> 32));
}

template <typename Half, Split kSplit>
int cross_argmin(hipStream_t s, const void* a, long long n_a, const void* b, long long n_b,
                 int n_sel, int K, int* index, float* value, unsigned long long* keys) {
    SplitBlock ba = carve_split(static_cast<char*>(const_cast<void*>(a)), n_a, K);
    SplitBlock bb = carve_split(static_cast<char*>(const_cast<void*>(b)), n_b, K);
    GramArgs<Half> g{};
    g.split_a = static_cast<const Half*>(ba.split);
    g.split_b = static_cast<const Half*>(bb.split);
    g.scale_a = ba.scale;
    g.scale_b = bb.scale;
    g.traces_a = ba.traces;
    g.traces_b = bb.traces;
    g.n_a = n_a;
    g.n_b = n_b;
    g.n_sel = n_sel;
    g.k_pad = K;
    g.tiles_b = (n_b + kMTile - 1) / kMTile;
    g.n_blocks = (n_a + kMTile - 1) / kMTile * g.tiles_b;
    g.out = nullptr;
    g.out_ld = 0;
    g.row_best = keys;
    if (g.n_blocks >= (1LL << 31)) {
        set_error("too many tiles (%lld)", g.n_blocks);
        return HR_ERR_UNSUPPORTED;
    }
    const unsigned blocks = (unsigned)((n_a + 255) / 256);
    hipLaunchKernelGGL(fill_keys_kernel, dim3(blocks), dim3(256), 0, s, keys, n_a);
    int rc = launch_status("fill_keys_kernel");
    if (rc) return rc;
    if ((rc = launch_pairs<Half, kSplit, true>(s, g))) return rc;
    hipLaunchKernelGGL(unpack_keys_kernel, dim3(blocks), dim3(256), 0, s, keys, n_a, index,
                       value);
    return launch_status("unpack_keys_kernel");
}

bool split_mode(int mode) {
    return mode == HR_PAIRS_F16X3 || mode == HR_PAIRS_BF16X3 || mode == HR_PAIRS_F16_EXACT ||
           mode == HR_PAIRS_BF16_EXACT;
}

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_rmsd_all_pairs_workspace_bytes(int64_t n_frames, int32_t n_sel, int32_t mode) {
    if (n_frames <= 0 || n_sel <= 0) return 256;
    const int K = k_pad_for(n_sel);
    if (mode == HR_PAIRS_FP32) return fp32_ws_bytes(n_frames, K);
    return carve_split(nullptr, n_frames, K).bytes;
}

int hr_rmsd_all_pairs(const float* xyz, int64_t n_frames, int32_t n_atoms_total,
                      const int32_t* atom_idx, int32_t n_sel, int32_t mode, float* out,
                      void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    HR_REQUIRE(mode == HR_PAIRS_FP32 || split_mode(mode), "unknown mode %d", mode);
    if (n_frames < 2) return HR_OK;
    HR_REQUIRE(xyz && out && workspace, "NULL pointer argument");
    HR_REQUIRE(workspace_bytes >= hr_rmsd_all_pairs_workspace_bytes(n_frames, n_sel, mode),
               "workspace too small: %zu < %zu", workspace_bytes,
               hr_rmsd_all_pairs_workspace_bytes(n_frames, n_sel, mode));
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const int K = k_pad_for(n_sel);
    const long long F = n_frames;
    switch (mode) {
        case HR_PAIRS_F16X3:
            return all_pairs_split<_Float16, Split::ThreeProduct>(s, workspace, xyz, F,
                                                                  n_atoms_total, atom_idx, n_sel,
                                                                  K, out);
        case HR_PAIRS_BF16X3:
            return all_pairs_split<__bf16, Split::ThreeProduct>(s, workspace, xyz, F,
                                                                n_atoms_total, atom_idx, n_sel, K,
                                                                out);
        case HR_PAIRS_F16_EXACT:
            return all_pairs_split<_Float16, Split::ExactHi>(s, workspace, xyz, F, n_atoms_total,
                                                             atom_idx, n_sel, K, out);
        case HR_PAIRS_BF16_EXACT:
            return all_pairs_split<__bf16, Split::ExactHi>(s, workspace, xyz, F, n_atoms_total,
                                                           atom_idx, n_sel, K, out);
        default:
            break;
    }
    char* ws = static_cast<char*>(workspace);
    double* traces = reinterpret_cast<double*>(ws);
    float* planes = reinterpret_cast<float*>(ws + align256((size_t)F * sizeof(double)));
    hipLaunchKernelGGL((pairs_prep_kernel<_Float16, Split::ThreeProduct>), dim3(prep_blocks_for(F)),
                       dim3(kPrepBlock), 0, s, xyz, F, n_atoms_total, atom_idx, n_sel, K, 0, planes,
                       traces, static_cast<_Float16*>(nullptr), nullptr, nullptr);
    int rc = launch_status("pairs_prep_kernel");
    if (rc) return rc;
    const long long T = (F + kF32Tile - 1) / kF32Tile;
    const long long n_tiles = T * (T + 1) / 2;
    HR_REQUIRE(n_tiles < (1LL << 31), "too many tiles (%lld)", n_tiles);
    hipLaunchKernelGGL(pairs_fp32_kernel, dim3((unsigned)n_tiles), dim3(kF32Block), 0, s, planes,
                       traces, F, n_sel, K, n_tiles, out);
    return launch_status("pairs_fp32_kernel");
}

size_t hr_split_bytes(int64_t n_frames, int32_t n_sel, int32_t mode) {
    if (n_frames <= 0 || n_sel <= 0 || !split_mode(mode)) return 0;
    return carve_split(nullptr, n_frames, k_pad_for(n_sel)).bytes;
}

int hr_split_prepare(const float* xyz, int64_t n_frames, int32_t n_atoms_total,
                     const int32_t* atom_idx, int32_t n_sel, int32_t mode, void* block,
                     size_t block_bytes, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    HR_REQUIRE(split_mode(mode), "mode %d is not a split mode", mode);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(xyz && block, "NULL pointer argument");
    HR_REQUIRE(block_bytes >= hr_split_bytes(n_frames, n_sel, mode), "block too small");
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const int K = k_pad_for(n_sel);
    SplitBlock b = carve_split(static_cast<char*>(block), n_frames, K);
    switch (mode) {
        case HR_PAIRS_F16X3:
            return launch_prep<_Float16, Split::ThreeProduct>(s, xyz, n_frames, n_atoms_total,
                                                              atom_idx, n_sel, K, b);
        case HR_PAIRS_BF16X3:
            return launch_prep<__bf16, Split::ThreeProduct>(s, xyz, n_frames, n_atoms_total,
                                                            atom_idx, n_sel, K, b);
        case HR_PAIRS_F16_EXACT:
            return launch_prep<_Float16, Split::ExactHi>(s, xyz, n_frames, n_atoms_total, atom_idx,
                                                         n_sel, K, b);
        default:
            return launch_prep<__bf16, Split::ExactHi>(s, xyz, n_frames, n_atoms_total, atom_idx,
                                                       n_sel, K, b);
    }
}

int hr_rmsd_cross(const void* block_a, int64_t n_a, const void* block_b, int64_t n_b,
                  int32_t n_sel, int32_t mode, float* out, int64_t out_ld, void* stream) {
    clear_error();
    HR_REQUIRE(n_a >= 0 && n_b >= 0 && n_sel > 0 && out_ld >= n_b,
               "invalid sizes n_a=%lld n_b=%lld n_sel=%d out_ld=%lld", (long long)n_a,
               (long long)n_b, n_sel, (long long)out_ld);
    HR_REQUIRE(split_mode(mode), "mode %d is not a split mode", mode);
    if (n_a == 0 || n_b == 0) return HR_OK;
    HR_REQUIRE(block_a && block_b && out, "NULL pointer argument");
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const int K = k_pad_for(n_sel);
    switch (mode) {
        case HR_PAIRS_F16X3:
            return cross_split<_Float16, Split::ThreeProduct>(s, block_a, n_a, block_b, n_b, n_sel,
                                                              K, out, out_ld);
        case HR_PAIRS_BF16X3:
            return cross_split<__bf16, Split::ThreeProduct>(s, block_a, n_a, block_b, n_b, n_sel,
                                                            K, out, out_ld);
        case HR_PAIRS_F16_EXACT:
            return cross_split<_Float16, Split::ExactHi>(s, block_a, n_a, block_b, n_b, n_sel, K,
                                                         out, out_ld);
        default:
            return cross_split<__bf16, Split::ExactHi>(s, block_a, n_a, block_b, n_b, n_sel, K,
                                                       out, out_ld);
    }
}

size_t hr_rmsd_cross_argmin_workspace_bytes(int64_t n_a) {
    return n_a > 0 ? (size_t)n_a * sizeof(unsigned long long) : 256;
}

int hr_rmsd_cross_argmin(const void* block_a, int64_t n_a, const void* block_b, int64_t n_b,
                         int32_t n_sel, int32_t mode, int32_t* index, float* value,
                         void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_a >= 0 && n_b >= 0 && n_sel > 0, "invalid sizes n_a=%lld n_b=%lld n_sel=%d",
               (long long)n_a, (long long)n_b, n_sel);
    HR_REQUIRE(split_mode(mode), "mode %d is not a split mode", mode);
    HR_REQUIRE(n_b > 0 || n_a == 0, "no columns to take the nearest of");
    if (n_a == 0) return HR_OK;
    HR_REQUIRE(n_b < (1LL << 31), "too many columns (%lld)", (long long)n_b);
    HR_REQUIRE(block_a && block_b && index && value && workspace, "NULL pointer argument");
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const int K = k_pad_for(n_sel);
    auto* keys = static_cast<unsigned long long*>(workspace);
    switch (mode) {
        case HR_PAIRS_F16X3:
            return cross_argmin<_Float16, Split::ThreeProduct>(s, block_a, n_a, block_b, n_b,
                                                               n_sel, K, index, value, keys);
        case HR_PAIRS_BF16X3:
            return cross_argmin<__bf16, Split::ThreeProduct>(s, block_a, n_a, block_b, n_b, n_sel,
                                                             K, index, value, keys);
        case HR_PAIRS_F16_EXACT:
            return cross_argmin<_Float16, Split::ExactHi>(s, block_a, n_a, block_b, n_b, n_sel, K,
                                                          index, value, keys);
        default:
            return cross_argmin<__bf16, Split::ExactHi>(s, block_a, n_a, block_b, n_b, n_sel, K,
                                                        index, value, keys);
    }
}

}  // extern "C"
