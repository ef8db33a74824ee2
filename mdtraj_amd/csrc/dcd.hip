// DCD ingestion (gfx950): CHARMM / NAMD / X-PLOR trajectory bytes, resident
// in HBM, turned into float32 (F, n_sel, 3) coordinates in HBM.
//
// Format (the reference's reader, formats/dcd/src/dcdplugin.c): Fortran
// records, each framed by its byte length before and after.  The header
// (read_dcdheader, :98-374) is the 84-byte control record ("CORD", NSET,
// ISTART, NSAVC, NAMNF, DELTA, and, for CHARMM files, the unit-cell and
// fourth-dimension flags and a version number), the title record, the atom
// count and, when NAMNF atoms are fixed, the 1-based free-atom list.  A frame
// (read_dcdstep, :481-623) is an optional unit-cell record (six doubles, A,
// cos gamma, B, cos beta, cos alpha, C), then the X, Y and Z records (every
// atom in frame 0; only the free atoms in later frames when atoms are fixed),
// then an optional fourth-dimension record that is skipped.  Byte order is
// the writer's, found from the first record marker.
//
// Frames after the first all have one size, so a frame's offset is
// arithmetic: the host parses the header (hr_dcd_scan) and the device turns
// the axis-major records into atom-major coordinates: an HBM-bound
// transposition, staged through LDS so that loads and stores are both
// coalesced.  The lane that handles atom 0 of a frame also checks every
// record marker of the frame (the reference's checks) and copies the unit
// cell; a frame whose markers are wrong gets a non-zero status word.
#include <string.h>

#include "common.h"

namespace hr {
namespace {

constexpr int kDcdBlock = 256;
constexpr int kPer = 4;                   // entries per lane
constexpr int kTile = kDcdBlock * kPer;   // entries per workgroup
constexpr int kCellBytes = 56;  // marker, six doubles, marker

struct DcdLayout {
    long long first;        // byte offset of frame 0
    long long first_bytes;  // bytes of frame 0
    long long frame_bytes;  // bytes of every later frame
    long long n_frames;     // frames in the file
    int n_atoms;
    int n_free;             // atoms stored in frames after the first
    int extra;              // unit-cell record present
    int four;               // fourth-dimension record present
    int reverse;            // the file's byte order is not the host's
};

__host__ __device__ __forceinline__ uint32_t swap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

__device__ __forceinline__ uint32_t word_at(const uint32_t* __restrict__ w, long long byte,
                                            int reverse) {
    const uint32_t v = w[byte >> 2];
    return reverse ? swap32(v) : v;
}

__device__ __forceinline__ long long frame_start(const DcdLayout& L, long long k) {
    return L.first + (k == 0 ? 0 : L.first_bytes + (k - 1) * L.frame_bytes);
}

// Record markers of one frame and its unit cell; returns the status word.
__device__ int check_frame(const uint32_t* __restrict__ w, const DcdLayout& L, long long off,
                           long long n, float* __restrict__ cell) {
    int st = HR_DCD_OK;
    float c[6] = {0.f, 90.f, 0.f, 90.f, 90.f, 0.f};  // the reference's values without a cell
    long long p = off;
    if (L.extra) {
        if (word_at(w, p, L.reverse) != 48u) {
            st = HR_DCD_BAD_RECORD;
        } else {
            for (int i = 0; i < 6; ++i) {
                const uint32_t a = w[(p + 4 + 8 * i) >> 2], b = w[(p + 8 + 8 * i) >> 2];
                // little-endian doubles: low word first; big-endian: both
                // words byte-swapped and the high word first
                const unsigned long long bits =
                    L.reverse ? ((unsigned long long)swap32(a) << 32) | swap32(b)
                              : ((unsigned long long)b << 32) | a;
                c[i] = (float)__longlong_as_double((long long)bits);
            }
        }
        p += kCellBytes;
    }
    const uint32_t len = (uint32_t)(4 * n);
    for (int r = 0; r < 3; ++r) {
        const long long q = p + r * (4 * n + 8);
        if (word_at(w, q, L.reverse) != len || word_at(w, q + 4 + 4 * n, L.reverse) != len)
            st = HR_DCD_BAD_RECORD;
    }
    // the fourth-dimension record is skipped by its leading length; this
    // layout needs it to match the other records
    if (L.four && word_at(w, p + 3 * (4 * n + 8), L.reverse) != len) st = HR_DCD_BAD_RECORD;
    if (cell)
        for (int i = 0; i < 6; ++i) cell[i] = c[i];
    return st;
}

// A workgroup owns kTile consecutive entries of the flattened (frame,
// selected atom) space, whose outputs are 3 kTile consecutive floats.  Each
// lane first works out the source words of its kPer entries and issues all
// 3 kPer loads back to back (invalid entries read word 0, which exists, and
// are replaced by NaN), then the entries go to LDS as x, y, z and the tile
// leaves in 3 kPer fully coalesced dword stores per lane.
__global__ __launch_bounds__(kDcdBlock) void dcd_decode_kernel(
    const uint32_t* __restrict__ words, long long n_bytes, DcdLayout L,
    const long long* __restrict__ frames, int n_frames, const int* __restrict__ sel, int n_sel,
    const int* __restrict__ free_slot, float scale, float* __restrict__ xyz,
    float* __restrict__ cell, int* __restrict__ status) {
    __shared__ float tile[3 * kTile];
    // n_frames * n_sel < 2^31 (the host splits larger reads)
    const unsigned total = (unsigned)n_frames * (unsigned)n_sel;
    const unsigned t0 = blockIdx.x * (unsigned)kTile;
    const long long first_end = L.first + L.first_bytes;
    // stages over the kPer entries, each issuing its loads together: frame
    // numbers, selected atoms, free-list slots, then the coordinates
    // (out-of-range entries use index 0 of each array, which exists)
    bool in[kPer], ok[kPer], full[kPer];
    unsigned f[kPer];
    int j[kPer], a[kPer], sl[kPer];
    long long k[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const unsigned g = t0 + m * kDcdBlock + threadIdx.x;
        in[m] = g < total;
        f[m] = in[m] ? g / (unsigned)n_sel : 0u;
        j[m] = in[m] ? (int)(g - f[m] * (unsigned)n_sel) : 0;
    }
#pragma unroll
    for (int m = 0; m < kPer; ++m) k[m] = frames[f[m]];
#pragma unroll
    for (int m = 0; m < kPer; ++m) a[m] = sel ? sel[j[m]] : j[m];
    long long off[kPer], c0[kPer], n[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        ok[m] = in[m] && k[m] >= 0 && k[m] < L.n_frames && first_end <= n_bytes;
        full[m] = k[m] == 0 || L.n_free == L.n_atoms;
        n[m] = full[m] ? L.n_atoms : L.n_free;
        off[m] = ok[m] ? frame_start(L, k[m]) : 0;
        c0[m] = off[m] + (L.extra ? kCellBytes : 0);  // X record marker
        ok[m] = ok[m] && c0[m] + (3 + L.four) * (4 * n[m] + 8) <= n_bytes;
    }
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const bool atom_ok = ok[m] && (unsigned)a[m] < (unsigned)L.n_atoms;
        // a free atom of a later frame sits at its slot in that frame's
        // records; a fixed atom keeps its coordinates from frame 0
        sl[m] = full[m] ? a[m] : free_slot[atom_ok ? a[m] : 0];
    }
    long long src[kPer], stride[kPer];
    bool valid[kPer];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        valid[m] = ok[m] && (unsigned)a[m] < (unsigned)L.n_atoms;
        if (sl[m] >= 0) {
            src[m] = c0[m] + 4 + 4LL * sl[m];
            stride[m] = 4 * n[m] + 8;
        } else {
            src[m] = L.first + (L.extra ? kCellBytes : 0) + 4 + 4LL * a[m];
            stride[m] = 4LL * L.n_atoms + 8;
        }
        if (!valid[m]) src[m] = stride[m] = 0;
    }
    uint32_t w[kPer][3];
#pragma unroll
    for (int m = 0; m < kPer; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) w[m][c] = words[(src[m] + c * stride[m]) >> 2];
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        if (!in[m] || j[m] != 0) continue;  // the frame's checks, by its atom-0 entry
        float* cf = cell ? cell + 6LL * f[m] : nullptr;
        status[f[m]] = ok[m] ? check_frame(words, L, off[m], n[m], cf) : HR_DCD_TRUNCATED;
    }
#pragma unroll
    for (int m = 0; m < kPer; ++m) {
        const int e = m * kDcdBlock + threadIdx.x;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const uint32_t u = L.reverse ? swap32(w[m][c]) : w[m][c];
            tile[3 * e + c] = valid[m] ? __fmul_rn(__uint_as_float(u), scale)
                                       : __int_as_float(0x7fc00000);
        }
    }
    __syncthreads();
    const unsigned left = total - t0;
    const unsigned n_out = 3u * (left < (unsigned)kTile ? left : (unsigned)kTile);
    float* out = xyz + 3LL * t0;
#pragma unroll
    for (int m = 0; m < 3 * kPer; ++m) {
        const unsigned o = m * kDcdBlock + threadIdx.x;
        if (o < n_out) out[o] = tile[o];
    }
}

}  // namespace
}  // namespace hr

using namespace hr;

namespace {

// Host-side header parse; mirrors read_dcdheader and the frame count of
// open_dcd_read (dcdplugin.c:98-374, 870-905).
int parse_header(const unsigned char* p, long long n_bytes, hr_dcd_header* h,
                 int32_t* free_indexes, int32_t max_free) {
    auto need = [&](long long end) { return end <= n_bytes; };
    HR_REQUIRE(need(8), "file too short for a DCD header (%lld bytes)", n_bytes);
    uint32_t w0, w1, cord;
    memcpy(&w0, p, 4);
    memcpy(&w1, p + 4, 4);
    memcpy(&cord, "CORD", 4);
    int reverse;
    if (w0 + w1 == 84u) {
        set_error("CHARMM -i8 DCD files (64-bit record markers) are not supported");
        return HR_ERR_UNSUPPORTED;
    } else if (w0 == 84u && w1 == cord) {
        reverse = 0;
    } else if (swap32(w0) + swap32(w1) == 84u) {
        set_error("CHARMM -i8 DCD files (64-bit record markers) are not supported");
        return HR_ERR_UNSUPPORTED;
    } else if (swap32(w0) == 84u && w1 == cord) {
        reverse = 1;
    } else {
        set_error("unrecognized DCD header (first words 0x%08x 0x%08x)", w0, w1);
        return HR_ERR_INVALID;
    }
    HR_REQUIRE(need(92), "truncated DCD header");
    const unsigned char* hdr = p + 8;
    auto raw = [&](int off) {  // flags are tested without byte swapping, as the reference
        int32_t v;
        memcpy(&v, hdr + off, 4);
        return v;
    };
    auto u32 = [&](long long off) {
        uint32_t v;
        memcpy(&v, p + off, 4);
        return reverse ? swap32(v) : v;
    };
    auto i32h = [&](int off) {
        uint32_t v;
        memcpy(&v, hdr + off, 4);
        return (int32_t)(reverse ? swap32(v) : v);
    };
    int flags = 0;
    if (raw(76) != 0) {
        flags = HR_DCD_CHARMM;
        if (raw(40) != 0) flags |= HR_DCD_EXTRA_BLOCK;
        if (raw(44) == 1) flags |= HR_DCD_FOUR_DIMS;
    }
    h->flags = flags;
    h->reverse_endian = reverse;
    h->n_set = i32h(0);
    h->istart = i32h(4);
    h->nsavc = i32h(8);
    const int32_t namnf = i32h(32);
    if (flags & HR_DCD_CHARMM) {
        uint32_t v;
        memcpy(&v, hdr + 36, 4);
        if (reverse) v = swap32(v);
        float d;
        memcpy(&d, &v, 4);
        h->delta = d;
    } else {
        uint32_t a, b;
        memcpy(&a, hdr + 36, 4);
        memcpy(&b, hdr + 40, 4);
        const unsigned long long bits = reverse ? ((unsigned long long)swap32(a) << 32) | swap32(b)
                                                : ((unsigned long long)b << 32) | a;
        memcpy(&h->delta, &bits, 8);
    }
    HR_REQUIRE(u32(88) == 84u, "bad end of the DCD control record");
    HR_REQUIRE(need(100), "truncated DCD header");
    HR_REQUIRE((u32(92) - 4u) % 80u == 0u, "bad DCD title record");
    int32_t ntitle = (int32_t)u32(96);
    HR_REQUIRE(ntitle >= 0, "negative DCD title count %d", ntitle);
    if (ntitle > 1000) ntitle = ntitle == 1095062083 ? 2 : 0;  // the reference's repairs
    long long pos = 100 + 80LL * ntitle + 4;
    HR_REQUIRE(need(pos + 12), "truncated DCD header");
    HR_REQUIRE(u32(pos) == 4u && u32(pos + 8) == 4u, "bad DCD atom-count record");
    const int32_t natoms = (int32_t)u32(pos + 4);
    pos += 12;
    if (natoms <= 0 || natoms > (1 << 30)) {
        set_error("unsupported DCD atom count %d", natoms);
        return HR_ERR_UNSUPPORTED;
    }
    HR_REQUIRE(namnf >= 0 && namnf < natoms, "bad DCD fixed-atom count %d of %d", namnf, natoms);
    const long long nfree = (long long)natoms - namnf;
    if (namnf > 0) {
        HR_REQUIRE(need(pos + 8 + 4 * nfree), "truncated DCD free-atom list");
        HR_REQUIRE(u32(pos) == 4 * nfree && u32(pos + 4 + 4 * nfree) == 4 * nfree,
                   "bad DCD free-atom record");
        for (long long i = 0; i < nfree; ++i) {
            const int32_t v = (int32_t)u32(pos + 4 + 4 * i);
            HR_REQUIRE(v >= 1 && v <= natoms, "DCD free-atom index %d out of range", v);
            if (free_indexes && i < max_free) free_indexes[i] = v;
        }
        pos += 8 + 4 * nfree;
    }
    const int ndims = (flags & HR_DCD_FOUR_DIMS) ? 4 : 3;
    const long long extra = (flags & HR_DCD_EXTRA_BLOCK) ? kCellBytes : 0;
    h->n_atoms = natoms;
    h->n_fixed = namnf;
    h->first_frame = pos;
    h->first_frame_bytes = ((long long)natoms + 2) * ndims * 4 + extra;
    h->frame_bytes = (nfree + 2) * ndims * 4 + extra;
    const long long trj = n_bytes - pos - h->first_frame_bytes;
    HR_REQUIRE(trj >= 0, "DCD file contains no frames");
    h->n_frames = trj / h->frame_bytes + 1;
    h->pad = 0;
    return HR_OK;
}

}  // namespace

extern "C" {

int hr_dcd_scan(const void* data, int64_t n_bytes, hr_dcd_header* header, int32_t* free_indexes,
                int32_t max_free) {
    clear_error();
    HR_REQUIRE(header, "NULL pointer argument");
    HR_REQUIRE(n_bytes >= 0 && (data || n_bytes == 0), "invalid buffer");
    HR_REQUIRE(max_free >= 0 && (free_indexes || max_free == 0), "invalid free-index buffer");
    memset(header, 0, sizeof(*header));
    return parse_header(static_cast<const unsigned char*>(data), n_bytes, header, free_indexes,
                        max_free);
}

int hr_dcd_decode(const void* data, int64_t n_bytes, const hr_dcd_header* header,
                  const int64_t* frames, int64_t n_frames, const int32_t* atom_indices,
                  int32_t n_sel, const int32_t* free_slot, float scale, float* xyz, float* cell,
                  int32_t* status, void* stream) {
    clear_error();
    HR_REQUIRE(header, "NULL pointer argument");
    const hr_dcd_header& h = *header;
    HR_REQUIRE(n_frames >= 0 && n_bytes >= 0 && n_sel > 0 && h.n_atoms > 0,
               "invalid sizes n_frames=%lld n_sel=%d n_atoms=%d", (long long)n_frames, n_sel,
               h.n_atoms);
    HR_REQUIRE(atom_indices || n_sel == h.n_atoms, "n_sel must equal n_atoms without atom_indices");
    HR_REQUIRE(h.n_fixed >= 0 && h.n_fixed < h.n_atoms, "invalid n_fixed %d", h.n_fixed);
    HR_REQUIRE(h.frame_bytes > 0 && h.first_frame >= 0 && h.n_frames > 0 &&
                   h.first_frame + h.first_frame_bytes <= n_bytes,
               "header does not describe these bytes");
    HR_REQUIRE(h.n_fixed == 0 || free_slot, "free_slot is required when atoms are fixed");
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(data && frames && xyz && status, "NULL pointer argument");
    HR_REQUIRE((reinterpret_cast<uintptr_t>(data) & 3) == 0, "data must be 4-byte aligned");
    DcdLayout L;
    L.first = h.first_frame;
    L.first_bytes = h.first_frame_bytes;
    L.frame_bytes = h.frame_bytes;
    L.n_frames = h.n_frames;
    L.n_atoms = h.n_atoms;
    L.n_free = h.n_atoms - h.n_fixed;
    L.extra = (h.flags & HR_DCD_EXTRA_BLOCK) ? 1 : 0;
    L.four = (h.flags & HR_DCD_FOUR_DIMS) ? 1 : 0;
    L.reverse = h.reverse_endian ? 1 : 0;
    // launches of at most 2^31 - 1 entries
    const long long per = std::max<long long>(1, ((1LL << 31) - 1) / n_sel);
    for (long long f0 = 0; f0 < n_frames; f0 += per) {
        const long long nf = std::min<long long>(per, n_frames - f0);
        const long long entries = nf * n_sel;
        const long long blocks = (entries + kTile - 1) / kTile;
        hipLaunchKernelGGL(dcd_decode_kernel, dim3((unsigned)blocks), dim3(kDcdBlock), 0,
                           (hipStream_t)stream, static_cast<const uint32_t*>(data),
                           (long long)n_bytes, L, reinterpret_cast<const long long*>(frames) + f0,
                           (int)nf, atom_indices, n_sel, free_slot, scale, xyz + f0 * n_sel * 3,
                           cell ? cell + 6 * f0 : nullptr, status + f0);
        const int rc = launch_status("dcd_decode_kernel");
        if (rc) return rc;
    }
    return HR_OK;
}

}  // extern "C"
