// Batched RMSD over many independent systems of uneven size in one graph
// launch (gfx950).
//
// A plan holds, in device memory, one descriptor per system.  Its hipGraph
// runs, on one stream, a reference prep kernel (one workgroup per system
// gathers its reference frame into axis-major planes and sums it in FP64)
// and then one scoring kernel, chosen by HR_TUNE_BATCH_SCHEDULE:
//   3 (default) batch_block_kernel: the frames of all systems form one
//     sequence, each frame costed as its bytes plus a fixed per-frame term
//     and each system as a fixed start-up term; workgroup g of G (three per
//     CU) owns the frames starting in its 1/G of the total cost.  For each
//     system in its range it stages the reference in LDS and its four waves
//     split the system's frames; frame sums stay in registers.
//   4 batch_cu_kernel: the same split with one workgroup per CU; the CU's
//     waves take a static share of each system's frames and then chunks from
//     an LDS counter.
//   1 batch_kernel: work units (system, up to 64 consecutive frames) sorted
//     by cost, largest first; wave w takes unit w first and then the next
//     unit from a ticket (vector atomic, requested while the current unit
//     runs), the reference streamed through the load ring with the frames.
//     A one-thread fill kernel zeroes the ticket first (no memset node).
//     Taking the first units without the ticket matters: when every wave of
//     the grid asks one counter at launch, the atomics to that address
//     serialise and the last waves start a third of a millisecond late
//     (100,000 x 2,000 in 33-frame units: 0.77 ms against 0.49 ms).  A plan
//     of this schedule must not be launched on two streams at once (the
//     ticket is shared).
//   2 batch_static_kernel: one contiguous cost-balanced range per wave, the
//     reference read from global memory with the frames.
// Every kernel's grid drains on every launch.
#include <algorithm>
#include <vector>

#include "frames.h"

struct hr_batch_plan {
    int device = -1;
    int n_systems = 0;
    long long n_units = 0;
    int frames_per_unit = 64;
    void* d_systems = nullptr;
    void* d_units = nullptr;
    unsigned* d_ticket = nullptr;
    long long* d_prefix = nullptr;  // static split: cost before system s, s = 0..n
    std::vector<long long> h_prefix;
    int* d_first = nullptr;         // block split: first system of each workgroup
    long long total_cost = 0;
    int schedule = 3;               // 1 ticket, 2 static split per wave, 3 per workgroup
    int lds_atoms = 0;              // block split: largest reference staged in LDS
    int max_atoms = 0;              // largest system with frames
    char* d_ref = nullptr;
    int blocks = 0;
    size_t lds = 0;
    long long total_frames = 0;
    double total_bytes = 0.0;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace hr {
namespace {

struct DevSystem {
    const float* xyz;
    long long n_frames;
    int n_atoms;
    int ref_frame;
    float* out;
    float* ref_sel;     // 3 n_atoms floats, axis-major
    double* ref_stats;  // sx, sy, sz, q
};

struct DevUnit {
    int system;
    int cnt;
    long long base;
};

__global__ void ticket_reset_kernel(unsigned* __restrict__ ticket) {
    if (threadIdx.x == 0) *ticket = 0u;
}

// One workgroup per system.  Its loads go out kPrepBatch at a time per thread
// (ref_prep_body): the node took 13.6 us per launch on the 512-system mix
// with one load in flight per thread, 10.1 us with eight; 1,024-thread
// workgroups were slower (12.5 us).
__global__ __launch_bounds__(kBlock) void batch_ref_prep_kernel(const DevSystem* __restrict__ sys) {
    const DevSystem s = sys[blockIdx.x];
    if (s.n_frames == 0) return;  // block-uniform: no barrier is skipped by part of it
    ref_prep_body(s.xyz + (long long)s.ref_frame * 3 * s.n_atoms, nullptr, s.n_atoms, 0,
                  s.ref_sel, s.ref_stats);
}

template <int kMinWaves>
__global__ __launch_bounds__(kBlock, kMinWaves) void batch_kernel(
    const DevSystem* __restrict__ sys, const DevUnit* __restrict__ units, long long n_units,
    unsigned* __restrict__ ticket) {
    extern __shared__ double smem_d[];
    constexpr int kNS = SumsPerFrame<Kind::Rmsd>::value;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    double* s_res = smem_d + wv * kWave * kNS;
    float bx[kRingB][kItemAtoms], by[kRingB][kItemAtoms], bz[kRingB][kItemAtoms];
    float qx[kRingB], qy[kRingB], qz[kRingB];
    // The first unit of every wave is its wave index, the rest come from the
    // ticket, which therefore counts from the grid's wave count: thousands of
    // waves taking their first unit from one atomic counter at launch would
    // start one after another as the atomics to that address serialise.
    const unsigned n_waves = gridDim.x * kWavesPerBlock;
    unsigned u = blockIdx.x * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    DevUnit unit{};
    DevSystem s{};
    if ((long long)u < n_units) {
        unit = units[u];
        s = sys[unit.system];
        prime_ring_batch(s.xyz, unit.base, unit.cnt, s.n_atoms, s.ref_sel, s.ref_sel + s.n_atoms,
                         s.ref_sel + 2 * s.n_atoms, lane, bx, by, bz, qx, qy, qz);
    }
    while ((long long)u < n_units) {
        // take the next ticket now; its latency hides behind this unit
        unsigned next = 0;
        if (lane == 0) next = atomicAdd(ticket, 1u);
        const int n = s.n_atoms;
        run_group_batch(s.xyz, unit.base, unit.cnt, n, s.ref_sel, s.ref_sel + n,
                        s.ref_sel + 2 * n, s_res, lane, bx, by, bz, qx, qy, qz);
        // start the next unit's loads before this one's QCP solves, so their
        // latency hides behind them; unconditionally (past the table's end
        // the last unit is loaded again and never used), since loads under a
        // branch are waited for at its join
        const unsigned nu = n_waves + __builtin_amdgcn_readfirstlane(__shfl(next, 0, kWave));
        const long long nuc = (long long)nu < n_units ? (long long)nu : n_units - 1;
        const DevUnit unit_n = units[nuc];
        const DevSystem s_n = sys[unit_n.system];
        prime_ring_batch(s_n.xyz, unit_n.base, unit_n.cnt, s_n.n_atoms, s_n.ref_sel,
                         s_n.ref_sel + s_n.n_atoms, s_n.ref_sel + 2 * s_n.n_atoms, lane, bx, by,
                         bz, qx, qy, qz);
        const GlobalDouble* st = global_ptr(s.ref_stats);
        const double rs[4] = {st[0], st[1], st[2], st[3]};
        finish_group<Kind::Rmsd>(unit.base, unit.cnt, (double)n, rs, (long long)s.ref_frame, s.out,
                                 nullptr, nullptr, s_res, lane);
        u = nu;
        unit = unit_n;
        s = s_n;
    }
}

// Cost of one frame in the static splits: its 12 N bytes plus the per-frame
// work that does not scale with N (row reduction, hand-off of the sums, QCP),
// which on MI355X costs about as much as streaming 600 more bytes.  Measured
// with the per-workgroup split on single systems of 10 - 4,400 atoms (the
// time difference between 0.3 GB and 0.9 GB of frames): ~2.07 ps per atom
// (5.8 TB/s) plus 45 - 130 ps per frame.
constexpr long long kFrameCost = 600;

__host__ __device__ __forceinline__ long long frame_cost(int n_atoms) {
    return 12LL * n_atoms + kFrameCost;
}

// Cost of taking up one more system in a workgroup's or wave's range: its
// descriptor, reference staging and first loads are latency the frames do
// not hide (about 4 us of streaming).  Without it a range of many small
// systems is balanced by bytes alone and finishes last.
constexpr long long kSystemCost = 24000;

// The frames [*fa, *fb) of a system whose cost starts at `start` (frame f at
// start + kSystemCost + f * frame_cost) that start in [c0, c1).
__device__ __forceinline__ void frames_in(long long start, long long n_frames, int n_atoms,
                                          long long c0, long long c1, long long* fa,
                                          long long* fb) {
    const long long pfc = frame_cost(n_atoms);
    const long long s0 = start + kSystemCost;
    *fa = c0 > s0 ? min(n_frames, (c0 - s0 + pfc - 1) / pfc) : 0;
    *fb = c1 > s0 ? min(n_frames, (c1 - s0 + pfc - 1) / pfc) : 0;
}

// Start of part g of `total` split into n near-equal parts.
__host__ __device__ __forceinline__ long long split_point(long long total, long long n,
                                                          long long g) {
    return total / n * g + total % n * g / n;
}

// One system's frames fa .. fb through the one-vs-all item loop, the row
// width chosen from its atom count, the reference read from global memory
// with the frame data (it is not staged in LDS: a wave's range can span
// several systems).
__device__ __forceinline__ void static_segment(const DevSystem& d, long long fa, long long fb,
                                               double* __restrict__ s_res, int lane) {
    const int n = d.n_atoms;
    const GlobalDouble* st = global_ptr(d.ref_stats);
    const double rs[4] = {st[0], st[1], st[2], st[3]};
#define HR_SEGMENT(RL)                                                                       \
    process_chunk<Kind::Rmsd, false, false, RL, true>(d.xyz, fa, fb, 3LL * n, 0, nullptr, n,  \
                                                      d.ref_sel, d.ref_sel + n,              \
                                                      d.ref_sel + 2 * n, rs,                 \
                                                      (long long)d.ref_frame, d.out, nullptr, \
                                                      nullptr, s_res, lane)
    switch (auto_row_lanes(n, false)) {
        case 4: HR_SEGMENT(4); break;
        case 8: HR_SEGMENT(8); break;
        case 16: HR_SEGMENT(16); break;
        default: HR_SEGMENT(32); break;
    }
#undef HR_SEGMENT
}

// Static split: the frames of all systems, in order, form one sequence whose
// frame (s, f) starts at cost prefix[s] + kSystemCost + f * frame_cost(N_s); wave w of W
// owns the frames starting in [C w / W, C (w + 1) / W), C the total cost -
// one contiguous range per wave, as the one-vs-all kernel splits a single
// trajectory, balanced by bytes plus per-frame work, with no atomics.
template <int kMinWaves>
__global__ __launch_bounds__(kBlock, kMinWaves) void batch_static_kernel(
    const DevSystem* __restrict__ sys, const long long* __restrict__ prefix, int n_systems,
    long long total) {
    extern __shared__ double smem_d[];
    constexpr int kNS = SumsPerFrame<Kind::Rmsd>::value;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double* s_res = smem_d + wv * kWave * kNS;
    const long long wave = (long long)blockIdx.x * kWavesPerBlock + wv;
    const long long n_waves = (long long)gridDim.x * kWavesPerBlock;
    const long long c0 = total / n_waves * wave + total % n_waves * wave / n_waves;
    const long long c1 = total / n_waves * (wave + 1) + total % n_waves * (wave + 1) / n_waves;
    // last system starting at or before c0
    int lo = 0, hi = n_systems;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (prefix[mid] <= c0) lo = mid;
        else hi = mid;
    }
    for (int si = lo; si < n_systems && prefix[si] < c1; ++si) {
        const DevSystem d = sys[si];
        long long fa, fb;
        frames_in(prefix[si], d.n_frames, d.n_atoms, c0, c1, &fa, &fb);
        if (fa < fb) static_segment(d, fa, fb, s_res, lane);
    }
}

// Block split (the default): as the static split, but per workgroup, so
// that the workgroup's four waves share one system at a time and its
// reference can sit in LDS, as in the one-vs-all kernel.  Workgroup g of G
// owns the frames starting in [C g / G, C (g + 1) / G); for each system in
// that range the workgroup copies the system's reference planes to LDS and
// its waves split the system's frames there into contiguous runs of whole
// frame sets.  A reference larger than lds_atoms atoms is staged in part
// (its first lds_atoms atoms) and read from global memory past that.  Frame
// sums stay in registers (process_chunk_regs), so LDS holds the reference
// alone.
template <int kMinWaves>
__global__ __launch_bounds__(kBlock, kMinWaves) void batch_block_kernel(
    const DevSystem* __restrict__ sys, const long long* __restrict__ prefix,
    const int* __restrict__ first, int n_systems, long long total, int lds_atoms) {
    extern __shared__ float s_ref[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const long long g = blockIdx.x, n_groups = gridDim.x;
    const long long c0 = split_point(total, n_groups, g);
    const long long c1 = split_point(total, n_groups, g + 1);
    const int lo = first[g];  // last system starting at or before c0 (host table)
    // every condition below depends only on the workgroup's range, so all
    // four waves take the same path through the barriers
    // the next system's descriptor is loaded while the current one runs
    DevSystem d = sys[lo];
    long long start = prefix[lo];
    for (int si = lo; si < n_systems && start < c1; ++si) {
        const int sn = min(si + 1, n_systems - 1);
        const DevSystem d_next = sys[sn];
        const long long start_next = si + 1 < n_systems ? prefix[sn] : c1;
        long long fa, fb;
        frames_in(start, d.n_frames, d.n_atoms, c0, c1, &fa, &fb);
        if (fa >= fb) {
            d = d_next;
            start = start_next;
            continue;
        }
        const int n = d.n_atoms;
        // the reference planes, or their first m atoms, in LDS (stride m),
        // copied once every wave's first frame loads are out
        const int m = min(n, lds_atoms);
        auto stage = [&]() {
            if (m > 0) {
                __syncthreads();  // the previous system's reference is no longer read
                stage_words(s_ref, d.ref_sel, m);
                stage_words(s_ref + m, d.ref_sel + n, m);
                stage_words(s_ref + 2 * m, d.ref_sel + 2 * n, m);
                __syncthreads();
            }
        };
        const GlobalDouble* st = global_ptr(d.ref_stats);
        const double rs[4] = {st[0], st[1], st[2], st[3]};
        const GlobalFloat* gx = global_ptr(d.ref_sel);
#define HR_BLOCK_SEGMENT(RL)                                                                  \
    do {                                                                                      \
        constexpr long long kRows = kWave / (RL);                                             \
        const long long n_sets = (fb - fa + kRows - 1) / kRows;                               \
        const long long wa = min(fb, fa + kRows * (wv * n_sets / kWavesPerBlock));            \
        const long long wb = min(fb, fa + kRows * ((wv + 1) * n_sets / kWavesPerBlock));      \
        if (m == n)                                                                           \
            process_chunk_regs<(RL), false>(d.xyz, wa, wb, n, s_ref, s_ref + m, s_ref + 2 * m, \
                                            m, gx, gx + n, gx + 2 * n, rs,                    \
                                            (long long)d.ref_frame, d.out, lane, stage);      \
        else                                                                                  \
            process_chunk_regs<(RL), true>(d.xyz, wa, wb, n, s_ref, s_ref + m, s_ref + 2 * m,  \
                                           m, gx, gx + n, gx + 2 * n, rs,                     \
                                           (long long)d.ref_frame, d.out, lane, stage);       \
    } while (0)
        switch (auto_row_lanes(n, false)) {
            case 4: HR_BLOCK_SEGMENT(4); break;
            case 8: HR_BLOCK_SEGMENT(8); break;
            case 16: HR_BLOCK_SEGMENT(16); break;
            default: HR_BLOCK_SEGMENT(32); break;
        }
#undef HR_BLOCK_SEGMENT
        d = d_next;
        start = start_next;
    }
}

// Per-CU split (schedule 4): as the block split, with one workgroup per CU
// holding all of its waves (kMinWaves per SIMD).  The workgroup stages each
// system's reference in the CU's LDS (up to ~13,600 atoms) and its waves
// share the system's frames: a static share of the first static_pct percent
// of the frame sets each, then chunks by an LDS ticket, so that the waves the
// SIMD issues less often take fewer frames (process_segment_cu).
template <int kMinWaves>
__global__ __launch_bounds__(CuGroup<kMinWaves>::kThreads, kMinWaves) void batch_cu_kernel(
    const DevSystem* __restrict__ sys, const long long* __restrict__ prefix,
    const int* __restrict__ first, int n_systems, long long total, int lds_atoms,
    int static_pct, int chunk_kib) {
    constexpr int kWaves = CuGroup<kMinWaves>::kWaves;
    constexpr int kThreads = CuGroup<kMinWaves>::kThreads;
    extern __shared__ float s_ref[];
    __shared__ unsigned s_ticket;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const long long g = blockIdx.x, n_groups = gridDim.x;
    const long long c0 = split_point(total, n_groups, g);
    const long long c1 = split_point(total, n_groups, g + 1);
    const int lo = first[g];
    DevSystem d = sys[lo];
    long long start = prefix[lo];
    for (int si = lo; si < n_systems && start < c1; ++si) {
        const int sn = min(si + 1, n_systems - 1);
        const DevSystem d_next = sys[sn];
        const long long start_next = si + 1 < n_systems ? prefix[sn] : c1;
        long long fa, fb;
        frames_in(start, d.n_frames, d.n_atoms, c0, c1, &fa, &fb);
        if (fa >= fb) {
            d = d_next;
            start = start_next;
            continue;
        }
        const int n = d.n_atoms;
        const int m = min(n, lds_atoms);
        auto stage = [&]() {
            __syncthreads();  // the previous system's reference and ticket are no longer used
            if (m > 0) {
                stage_words<kThreads>(s_ref, d.ref_sel, m);
                stage_words<kThreads>(s_ref + m, d.ref_sel + n, m);
                stage_words<kThreads>(s_ref + 2 * m, d.ref_sel + 2 * n, m);
            }
            if (threadIdx.x == 0) s_ticket = 0u;
            __syncthreads();
        };
        const GlobalDouble* st = global_ptr(d.ref_stats);
        const double rs[4] = {st[0], st[1], st[2], st[3]};
        const GlobalFloat* gx = global_ptr(d.ref_sel);
#define HR_CU_SEGMENT(RL)                                                                       \
    do {                                                                                        \
        constexpr long long kRows = kWave / (RL);                                               \
        const long long n_sets = (fb - fa + kRows - 1) / kRows;                                 \
        const long long n_static = n_sets * static_pct / 100;                                   \
        const long long sa = min(fb, fa + kRows * (wv * n_static / kWaves));                    \
        const long long sb = min(fb, fa + kRows * ((wv + 1) * n_static / kWaves));              \
        const long long da = min(fb, fa + kRows * n_static);                                    \
        long long ch = ((long long)chunk_kib * 1024 + 12LL * n - 1) / (12LL * n);               \
        ch = max(kRows, min((long long)kWave, (ch + kRows - 1) / kRows * kRows));               \
        if (m == n)                                                                             \
            process_segment_cu<(RL), false>(d.xyz, sa, sb, da, fb, ch, &s_ticket, n, s_ref,    \
                                            s_ref + m, s_ref + 2 * m, m, gx, gx + n,            \
                                            gx + 2 * n, rs, (long long)d.ref_frame, d.out,      \
                                            lane, stage);                                       \
        else                                                                                    \
            process_segment_cu<(RL), true>(d.xyz, sa, sb, da, fb, ch, &s_ticket, n, s_ref,     \
                                           s_ref + m, s_ref + 2 * m, m, gx, gx + n, gx + 2 * n, \
                                           rs, (long long)d.ref_frame, d.out, lane, stage);    \
    } while (0)
        switch (auto_row_lanes(n, false)) {
            case 4: HR_CU_SEGMENT(4); break;
            case 8: HR_CU_SEGMENT(8); break;
            case 16: HR_CU_SEGMENT(16); break;
            default: HR_CU_SEGMENT(32); break;
        }
#undef HR_CU_SEGMENT
        d = d_next;
        start = start_next;
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

void release(hr_batch_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->capture) (void)hipStreamDestroy(p->capture);
    if (p->d_systems) (void)hipFree(p->d_systems);
    if (p->d_units) (void)hipFree(p->d_units);
    if (p->d_ticket) (void)hipFree(p->d_ticket);
    if (p->d_prefix) (void)hipFree(p->d_prefix);
    if (p->d_first) (void)hipFree(p->d_first);
    if (p->d_ref) (void)hipFree(p->d_ref);
    delete p;
}

template <int kMinWaves>
int capture_plan(hr_batch_plan* p) {
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    const int sched = p->schedule;
    const void* k = sched == 4   ? reinterpret_cast<const void*>(batch_cu_kernel<kMinWaves>)
                    : sched == 3 ? reinterpret_cast<const void*>(batch_block_kernel<kMinWaves>)
                    : sched == 2 ? reinterpret_cast<const void*>(batch_static_kernel<kMinWaves>)
                                 : reinterpret_cast<const void*>(batch_kernel<kMinWaves>);
    const int threads = sched == 4 ? CuGroup<kMinWaves>::kThreads : kBlock;
    size_t lds = (size_t)kWavesPerBlock * kWave * SumsPerFrame<Kind::Rmsd>::value *
                 sizeof(double);
    if (sched == 4) {
        // one workgroup per CU: the CU's LDS (less the static ticket)
        const size_t cap = std::min(d->lds_per_cu, d->lds_per_block) / 1024 * 1024 - 1024;
        p->lds_atoms = (int)std::min<long long>(p->max_atoms, (long long)cap / 12 / 4 * 4);
        lds = (size_t)12 * p->lds_atoms;
        static bool attr_set = false;
        if (!attr_set) {
            HR_CHECK_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)cap));
            attr_set = true;
        }
    } else if (sched == 3) {
        // the largest reference, up to what leaves room for kMinWaves
        // workgroups per CU.  A workgroup's LDS is allocated in granules
        // coarser than the occupancy query assumes: on MI355X 54,240 B passes
        // the query for three per CU, yet only two are resident and a third
        // of the grid starts when the first ones end (~200 us late on the
        // 512-system mix: 0.507 ms against 0.449 ms); 53,232 B (4,436
        // atoms) keeps all three resident.  The share is rounded down to
        // 2 KiB.
        const size_t per_group =
            std::min(d->lds_per_cu / kMinWaves, d->lds_per_block) / 2048 * 2048;
        const long long fit = (long long)per_group / 12 / 4 * 4;
        p->lds_atoms = (int)std::min<long long>(p->max_atoms, fit);
        for (;;) {
            int per_cu = 0;
            lds = (size_t)12 * p->lds_atoms;
            HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, lds));
            if (per_cu >= kMinWaves || p->lds_atoms == 0) break;
            p->lds_atoms = std::max(0, p->lds_atoms - 64);
        }
    }
    int per_cu = 0;
    HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, lds));
    long long blocks = (long long)(per_cu < 1 ? 1 : per_cu) * d->cu_count;
    // ticket: no more waves than units; static splits: no more than frames
    const long long work = sched == 1 ? p->n_units : p->total_frames;
    const long long by_work = (work + threads / kWave - 1) / (threads / kWave);
    if (blocks > by_work) blocks = by_work;
    if (blocks < 1) blocks = 1;
    p->blocks = (int)blocks;
    p->lds = lds;
    if (sched == 3 || sched == 4) {
        // each workgroup's first system: the last one starting at or before
        // the workgroup's first cost
        std::vector<int> first(p->blocks);
        for (int g = 0; g < p->blocks; ++g) {
            const long long c0 = split_point(p->total_cost, p->blocks, g);
            const auto it = std::upper_bound(p->h_prefix.begin(),
                                             p->h_prefix.begin() + p->n_systems, c0);
            first[g] = std::max(0, (int)(it - p->h_prefix.begin()) - 1);
        }
        HR_CHECK_HIP(hipMalloc(&p->d_first, sizeof(int) * first.size()));
        HR_CHECK_HIP(hipMemcpy(p->d_first, first.data(), sizeof(int) * first.size(),
                               hipMemcpyHostToDevice));
    }
    hipStream_t s = p->capture;
    HR_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (sched == 1) hipLaunchKernelGGL(ticket_reset_kernel, dim3(1), dim3(64), 0, s, p->d_ticket);
    hipLaunchKernelGGL(batch_ref_prep_kernel, dim3(p->n_systems), dim3(kBlock), 0, s,
                       static_cast<const DevSystem*>(p->d_systems));
    if (sched == 4)
        hipLaunchKernelGGL(batch_cu_kernel<kMinWaves>, dim3(p->blocks), dim3(threads), lds, s,
                           static_cast<const DevSystem*>(p->d_systems), p->d_prefix, p->d_first,
                           p->n_systems, p->total_cost, p->lds_atoms,
                           tuning(HR_TUNE_FRAMES_STATIC_PCT) ? tuning(HR_TUNE_FRAMES_STATIC_PCT)
                                                             : kDefaultStaticPct,
                           tuning(HR_TUNE_FRAMES_CHUNK_KIB) ? tuning(HR_TUNE_FRAMES_CHUNK_KIB)
                                                            : kDefaultChunkKiB);
    else if (sched == 3)
        hipLaunchKernelGGL(batch_block_kernel<kMinWaves>, dim3(p->blocks), dim3(kBlock), lds, s,
                           static_cast<const DevSystem*>(p->d_systems), p->d_prefix, p->d_first,
                           p->n_systems, p->total_cost, p->lds_atoms);
    else if (sched == 2)
        hipLaunchKernelGGL(batch_static_kernel<kMinWaves>, dim3(p->blocks), dim3(kBlock), lds, s,
                           static_cast<const DevSystem*>(p->d_systems), p->d_prefix,
                           p->n_systems, p->total_cost);
    else
        hipLaunchKernelGGL(batch_kernel<kMinWaves>, dim3(p->blocks), dim3(kBlock), lds, s,
                           static_cast<const DevSystem*>(p->d_systems),
                           static_cast<const DevUnit*>(p->d_units), p->n_units, p->d_ticket);
    hipError_t launch = hipGetLastError();
    hipError_t end = hipStreamEndCapture(s, &p->graph);
    if (launch != hipSuccess || end != hipSuccess) {
        set_error("batch graph capture failed: %s / %s", hipGetErrorString(launch),
                  hipGetErrorString(end));
        return HR_ERR_HIP;
    }
    HR_CHECK_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return HR_OK;
}

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

int hr_batch_plan_create(const hr_system_desc* systems, int32_t n_systems,
                         const hr_batch_shape* shape, hr_batch_plan** plan_out) {
    clear_error();
    HR_REQUIRE(plan_out != nullptr, "plan_out is NULL");
    *plan_out = nullptr;
    HR_REQUIRE(systems != nullptr && n_systems > 0, "no systems");
    int frames_per_unit = shape ? shape->frames_per_unit : 0;
    const int unit_kib = shape ? shape->unit_kib : 0;
    int waves = shape ? shape->waves_per_simd : 0;
    if (frames_per_unit == 0) frames_per_unit = 64;
    if (waves == 0) waves = 3;
    HR_REQUIRE(frames_per_unit >= 1 && frames_per_unit <= kWave,
               "frames_per_unit must be in [1, 64] (got %d)", frames_per_unit);
    HR_REQUIRE(unit_kib >= 0, "unit_kib must be >= 0 (got %d)", unit_kib);
    HR_REQUIRE(waves == 3 || waves == 4, "waves_per_simd must be 3 or 4 (got %d)", waves);
    for (int i = 0; i < n_systems; ++i) {
        const hr_system_desc& s = systems[i];
        HR_REQUIRE(s.n_frames >= 0 && s.n_atoms > 0, "system %d: invalid sizes", i);
        HR_REQUIRE(s.n_frames == 0 || (s.ref_frame >= 0 && s.ref_frame < s.n_frames),
                   "system %d: ref_frame %d out of range", i, s.ref_frame);
        HR_REQUIRE(s.n_frames == 0 || (s.xyz && s.out), "system %d: NULL pointer", i);
    }
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;

    hr_batch_plan* p = new hr_batch_plan;
    p->device = d->device;
    p->n_systems = n_systems;
    p->frames_per_unit = frames_per_unit;
    const int sched = tuning(HR_TUNE_BATCH_SCHEDULE);
    p->schedule = sched == 0 ? 3 : sched;
    // static split: cost of the frames before each system
    std::vector<long long> prefix(n_systems + 1, 0);
    for (int i = 0; i < n_systems; ++i)
        prefix[i + 1] = prefix[i] + (systems[i].n_frames > 0 ? kSystemCost : 0) +
                        systems[i].n_frames * frame_cost(systems[i].n_atoms);
    p->total_cost = prefix[n_systems];
    p->h_prefix = prefix;

    // reference workspace: per system planes (3 N floats) then 4 doubles
    std::vector<size_t> ref_off(n_systems);
    size_t ref_bytes = 0;
    for (int i = 0; i < n_systems; ++i) {
        ref_off[i] = ref_bytes;
        ref_bytes += align256((size_t)systems[i].n_atoms * 3 * sizeof(float)) + 256;
    }
    // work units, largest first (LPT order keeps the tail short)
    std::vector<DevUnit> units;
    for (int i = 0; i < n_systems; ++i) {
        int fpu = frames_per_unit;
        if (unit_kib > 0) {
            // frames of 12 N bytes that fit the budget, as a multiple of 4 (>= 4)
            const long long fit = (long long)unit_kib * 1024 / (12LL * systems[i].n_atoms);
            const int cap = (int)std::max<long long>(4, fit / 4 * 4);
            fpu = std::min(fpu, cap);
        }
        for (long long b = 0; b < systems[i].n_frames; b += fpu) {
            const long long c = std::min<long long>(fpu, systems[i].n_frames - b);
            units.push_back(DevUnit{i, (int)c, b});
        }
        p->total_frames += systems[i].n_frames;
        if (systems[i].n_frames > 0) p->max_atoms = std::max(p->max_atoms, systems[i].n_atoms);
        p->total_bytes += 12.0 * (double)systems[i].n_frames * systems[i].n_atoms;
    }
    std::stable_sort(units.begin(), units.end(), [&](const DevUnit& a, const DevUnit& b) {
        return (long long)a.cnt * systems[a.system].n_atoms >
               (long long)b.cnt * systems[b.system].n_atoms;
    });
    p->n_units = (long long)units.size();

    int rc = HR_OK;
    hipError_t e = hipSuccess;
#define HR_PLAN_TRY(expr)                                                                     \
    do {                                                                                      \
        e = (expr);                                                                           \
        if (e != hipSuccess) {                                                                \
            set_error("%s failed: %s", #expr, hipGetErrorString(e));                           \
            release(p);                                                                       \
            return HR_ERR_HIP;                                                                \
        }                                                                                     \
    } while (0)
    HR_PLAN_TRY(hipMalloc(&p->d_ref, ref_bytes > 0 ? ref_bytes : 256));
    std::vector<DevSystem> ds(n_systems);
    for (int i = 0; i < n_systems; ++i) {
        const hr_system_desc& s = systems[i];
        char* base = p->d_ref + ref_off[i];
        ds[i] = DevSystem{s.xyz, s.n_frames, s.n_atoms, s.ref_frame, s.out,
                          reinterpret_cast<float*>(base),
                          reinterpret_cast<double*>(
                              base + align256((size_t)s.n_atoms * 3 * sizeof(float)))};
    }
    HR_PLAN_TRY(hipMalloc(&p->d_systems, sizeof(DevSystem) * n_systems));
    HR_PLAN_TRY(hipMalloc(&p->d_units, sizeof(DevUnit) * (units.empty() ? 1 : units.size())));
    HR_PLAN_TRY(hipMalloc(&p->d_ticket, 256));
    HR_PLAN_TRY(hipMalloc(&p->d_prefix, sizeof(long long) * (n_systems + 1)));
    HR_PLAN_TRY(hipMemcpy(p->d_prefix, prefix.data(), sizeof(long long) * (n_systems + 1),
                          hipMemcpyHostToDevice));
    HR_PLAN_TRY(hipMemcpy(p->d_systems, ds.data(), sizeof(DevSystem) * n_systems,
                          hipMemcpyHostToDevice));
    if (!units.empty())
        HR_PLAN_TRY(hipMemcpy(p->d_units, units.data(), sizeof(DevUnit) * units.size(),
                              hipMemcpyHostToDevice));
    HR_PLAN_TRY(hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking));
#undef HR_PLAN_TRY
    rc = waves == 4 ? capture_plan<4>(p) : capture_plan<3>(p);
    if (rc) {
        release(p);
        return rc;
    }
    *plan_out = p;
    return HR_OK;
}

int hr_batch_plan_launch(hr_batch_plan* plan, void* stream) {
    clear_error();
    HR_REQUIRE(plan != nullptr && plan->exec != nullptr, "invalid plan");
    HR_CHECK_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    return HR_OK;
}

int hr_batch_plan_destroy(hr_batch_plan* plan) {
    clear_error();
    release(plan);
    return HR_OK;
}

int64_t hr_batch_plan_units(const hr_batch_plan* plan) { return plan ? plan->n_units : -1; }

}  // extern "C"
