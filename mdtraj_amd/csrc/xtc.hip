// XTC ingestion (gfx950): GROMACS-compressed trajectory bytes, resident in
// HBM, decoded into float32 (F, N, 3) coordinates in HBM, so serving pools
// and one-vs-all queries read a file without a host-side decode.
//
// Format (the reference's reader: formats/xtc/src/xdrfile_xtc.c:41-90 for
// the frame header and box, formats/xtc/src/xdrfile.c:778-1000 for the
// coordinates): big-endian 32-bit words; a frame is magic 1995, natoms,
// step, time, nine box floats, natoms again, then either 3 natoms raw
// floats (natoms <= 9) or precision, minint[3], maxint[3], smallidx, a byte
// count and that many bytes of bit stream (padded to a word).  The stream
// holds, per "large" atom, its integer coordinates less minint (one
// mixed-radix number of sizeofints(maxint - minint + 1) bits, or three
// fields when a range exceeds 2^24), a flag bit, an optional 5-bit run code
// (run length in floats, plus a -1/0/+1 change of smallidx), and the run's
// small atoms as deltas (one mixed-radix number of smallidx bits each).
//
// Frames are independent, so the host walks the headers once (hr_xtc_scan,
// the frame index) and the device decodes one frame per lane: the bit
// stream inside a frame is sequential.  Every read is bounded by the frame's
// own bytes and every write by its atom count; a malformed frame ends with a
// non-zero status word (HR_XTC_*) instead of an out-of-range access.
#include <string.h>

#include "common.h"

namespace hr {
namespace {

constexpr int kXtcMagic = 1995;
constexpr int kFirstIdx = 9;
// Sizes of the small-delta digits by smallidx (the format's table: the
// value at index i needs i bits for three digits).
__constant__ int kMagicInts[] = {
    0,       0,       0,       0,       0,       0,       0,       0,       0,       8,
    10,      12,      16,      20,      25,      32,      40,      50,      64,      80,
    101,     128,     161,     203,     256,     322,     406,     512,     645,     812,
    1024,    1290,    1625,    2048,    2580,    3250,    4096,    5060,    6501,    8192,
    10321,   13003,   16384,   20642,   26007,   32768,   41285,   52015,   65536,   82570,
    104031,  131072,  165140,  208063,  262144,  330280,  416127,  524287,  660561,  832255,
    1048576, 1321122, 1664510, 2097152, 2642245, 3329021, 4194304, 5284491, 6658042, 8388607,
    10568983, 13316085, 16777216};
constexpr int kLastIdx = 73;  // entries in kMagicInts
static_assert(sizeof(kMagicInts) / sizeof(kMagicInts[0]) == kLastIdx, "table size");

__host__ __device__ __forceinline__ uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

__device__ __forceinline__ uint32_t be32_word(const uint32_t* w) {
    return __builtin_bswap32(*w);
}

// Bit length of v (bits needed for 0 .. v - 1 ... v itself: the format's
// sizeofint, and sizeofints of a product).
__device__ __forceinline__ int bit_length64(uint64_t v) { return v ? 64 - __clzll(v) : 0; }

// floor(v / d) and v mod d for v < 2^50 from a double reciprocal: the
// estimate is within one of the quotient (relative error below 2^-51 on a
// quotient below 2^50), and one integer step corrects it.
__device__ __forceinline__ uint64_t div_by_recip(uint64_t v, uint32_t d, double inv_d,
                                                 uint32_t& rem) {
    uint64_t q = (uint64_t)((double)v * inv_d);
    long long r = (long long)v - (long long)(q * d);
    if (r < 0) {
        q -= 1;
        r += d;
    } else if (r >= (long long)d) {
        q += 1;
        r -= d;
    }
    rem = (uint32_t)r;
    return q;
}

// MSB-first reader over one frame's bit stream (whole big-endian words;
// the stream is word-aligned and padded to a word in the file).
// The words come in blocks of kBlockWords registers loaded together: a lane
// reads its frame's words in order, and with few waves per SIMD (one lane
// per frame) nothing hides a load's latency, so each block costs one
// latency where word-by-word loads cost one per word.
constexpr int kBlockWords = 16;

struct BitReader {
    const uint32_t* w;
    int n_words;
    int consumed;   // words moved into buf
    int left;       // words of blk not yet moved into buf (blk[0] is next)
    uint32_t blk[kBlockWords];
    uint64_t buf;   // the low `have` bits are unread
    int have;
    bool bad;

    __device__ __forceinline__ void start(const uint32_t* words, int count) {
        w = words;
        n_words = count;
        consumed = 0;
        left = 0;
        buf = 0;
        have = 0;
        bad = false;
    }

    __device__ __forceinline__ void refill() {
        const int avail = n_words - consumed;  // > 0: word `consumed` exists
#pragma unroll
        for (int i = 0; i < kBlockWords; ++i) {
            const uint32_t x = w[consumed + (i < avail ? i : 0)];
            blk[i] = i < avail ? x : 0u;
        }
        left = avail < kBlockWords ? avail : kBlockWords;
    }

    __device__ __forceinline__ uint32_t take(int n) {  // 0 <= n <= 32
        if (n == 0) return 0;
        if (have < n) {
            if (consumed >= n_words) {
                bad = true;
                return 0;
            }
            if (left == 0) refill();
            buf = (buf << 32) | __builtin_bswap32(blk[0]);
#pragma unroll
            for (int i = 0; i < kBlockWords - 1; ++i) blk[i] = blk[i + 1];
            --left;
            ++consumed;
            have += 32;
        }
        have -= n;
        return (uint32_t)(buf >> have) & (uint32_t)(((uint64_t)1 << n) - 1);
    }

    // Three digits (x, y, z) of one mixed-radix number of nbits bits whose
    // 8-bit chunks come least significant first: z = V mod sz, y = (V / sz)
    // mod sy, x = V / (sz sy) (low 32 bits).
    // inv_sz, inv_sy: 1 / sz and 1 / sy in double when the divisors are the
    // frame's large-atom sizes (0 otherwise).
    __device__ __forceinline__ void ints3(int nbits, uint32_t sx, uint32_t sy, uint32_t sz,
                                          double inv_sz, double inv_sy, int32_t (&out)[3]) {
        (void)sx;
        if (nbits <= 64) {
            // the chunks are a = (nbits - 1) / 8 whole bytes, then the last
            // b = nbits - 8 a bits: read the field at once (most significant
            // bit first) and reverse the byte order of its whole bytes
            uint64_t w;
            if (nbits > 32) {
                const uint64_t hi = take(nbits - 32);
                w = (hi << 32) | take(32);
            } else {
                w = take(nbits);
            }
            const int a = (nbits - 1) >> 3, b = nbits - 8 * a;
            uint64_t v = w & ((1ull << b) - 1);
            if (a) v = (v << (8 * a)) | (__builtin_bswap64(w >> b) >> (64 - 8 * a));
            if (nbits <= 32) {  // the common case for small atoms: 32-bit division
                const uint32_t v32 = (uint32_t)v;
                const uint32_t q = v32 / sz;
                out[2] = (int32_t)(v32 - q * sz);
                const uint32_t q2 = q / sy;
                out[1] = (int32_t)(q - q2 * sy);
                out[0] = (int32_t)q2;
                return;
            }
            uint64_t q, q2;
            uint32_t r, r2;
            if (inv_sz > 0.0 && v < (1ull << 50)) {  // the large atoms of a frame
                q = div_by_recip(v, sz, inv_sz, r);
                q2 = div_by_recip(q, sy, inv_sy, r2);
            } else {
                q = v / sz;
                r = (uint32_t)(v - q * sz);
                q2 = q / sy;
                r2 = (uint32_t)(q - q2 * sy);
            }
            out[2] = (int32_t)r;
            out[1] = (int32_t)r2;
            out[0] = (int32_t)(uint32_t)q2;
            return;
        }
        // more than 64 bits (three sizes near 2^24): long division over the
        // byte digits, most significant first
        uint32_t bytes[12];
        int nb = 0;
        while (nbits > 8 && nb < 12) {
            bytes[nb++] = take(8);
            nbits -= 8;
        }
        if (nbits > 0 && nb < 12) bytes[nb++] = take(nbits);
        const uint32_t div[2] = {sz, sy};
        for (int d = 0; d < 2; ++d) {
            uint32_t rem = 0;
            for (int j = nb - 1; j >= 0; --j) {
                const uint32_t num = (rem << 8) | bytes[j];
                bytes[j] = num / div[d];
                rem = num - bytes[j] * div[d];
            }
            out[2 - d] = (int32_t)rem;
        }
        out[0] = (int32_t)(bytes[0] | (bytes[1] << 8) | (bytes[2] << 16) | (bytes[3] << 24));
    }
};

// A lane's decoded floats go out through a ring of two 128-byte lines in
// LDS.  Once per decoded run the lines that are complete are written: a line
// wholly inside the frame's output as eight 16-byte stores, so HBM sees full
// lines; the two lines a frame shares with its neighbours float by float.
// Written directly, the 64 lanes of a wave each dribbled 4-byte stores into 64
// different lines, which were evicted part-written and merged by
// read-modify-write: 2.3 times the algorithmic traffic.  Positions are
// absolute float addresses (byte address / 4), so ring lines are 128-byte
// lines of memory whatever the array's offset.
//
// Capacity: after a flush at most 31 floats of the current line are pending
// and one run adds at most 11 atoms (a 5-bit run length of at most 30 small
// atoms in threes, plus the large atom), 33 floats: 64 floats, the ring.
constexpr int kLineFloats = 32;
constexpr int kRingFloats = 2 * kLineFloats;
constexpr int kRingStride = kRingFloats + 4;  // 16-byte aligned rows, spread over the banks
constexpr int kMaxFloatsPerRun = 3 * (1 + 30 / 3);
static_assert(kLineFloats - 1 + kMaxFloatsPerRun <= kRingFloats, "ring too small for one run");

struct LineWriter {
    float* __restrict__ ring;  // this lane's ring in LDS
    uint64_t pos;              // absolute float address of the next float
    uint64_t line;             // first line (absolute float address) not yet written
    uint64_t beg, end;         // the frame's floats

    __device__ __forceinline__ void start(float* r, uint64_t first, uint64_t n) {
        ring = r;
        pos = beg = first;
        end = first + n;
        line = first & ~(uint64_t)(kLineFloats - 1);
    }

    __device__ __forceinline__ void put(float v) {
        ring[pos & (kRingFloats - 1)] = v;
        ++pos;
    }

    // the format's rounding: the integer to float, times float(1 / precision)
    __device__ __forceinline__ void put_atom(const int32_t (&c)[3], float inv) {
        put(__fmul_rn((float)c[0], inv));
        put(__fmul_rn((float)c[1], inv));
        put(__fmul_rn((float)c[2], inv));
    }

    // writes the line at `line` holding the floats below `upto`
    __device__ __forceinline__ void write_line(uint64_t upto) {
        const float* src = ring + (line & (kRingFloats - 1));
        if (line >= beg && line + kLineFloats <= upto) {
            typedef float v4f __attribute__((ext_vector_type(4)));
            const v4f* s4 = reinterpret_cast<const v4f*>(src);
            auto* dst = global_out(reinterpret_cast<v4f*>(line << 2));
#pragma unroll
            for (int k = 0; k < kLineFloats / 4; ++k) dst[k] = s4[k];
        } else {
            for (int k = 0; k < kLineFloats; ++k) {
                const uint64_t p = line + k;
                if (p >= beg && p < upto) *global_out(reinterpret_cast<float*>(p << 2)) = src[k];
            }
        }
    }

    // writes every complete line
    __device__ __forceinline__ void flush() {
        while (line + kLineFloats <= pos) {
            write_line(pos);
            line += kLineFloats;
        }
    }

    // writes what is left: the frame's last, partial line
    __device__ __forceinline__ void finish() {
        flush();
        if (line < pos) write_line(pos);
    }
};

// One frame per lane.  offsets[f] .. offsets[f + 1] (or n_bytes) bound the
// frame's bytes; the host has checked magic and sizes when it built them.
__global__ __launch_bounds__(kWave) void xtc_decode_kernel(
    const unsigned char* __restrict__ data, long long n_bytes,
    const long long* __restrict__ offsets, long long n_frames, int n_atoms,
    float* __restrict__ xyz, float* __restrict__ time, int* __restrict__ step,
    float* __restrict__ box, float* __restrict__ prec, int* __restrict__ status) {
    __shared__ __align__(16) float s_ring[kWave * kRingStride];
    const long long f = (long long)blockIdx.x * kWave + threadIdx.x;
    if (f >= n_frames) return;
    const long long beg = offsets[f];
    const long long end = f + 1 < n_frames ? offsets[f + 1] : n_bytes;
    int st = HR_XTC_OK;
    const unsigned char* p = data + beg;
    float* out = xyz + (long long)f * 3 * n_atoms;
    if (beg < 0 || end > n_bytes || end - beg < 56 || (beg & 3)) {
        status[f] = HR_XTC_TRUNCATED;
        return;
    }
    const uint32_t* pw = reinterpret_cast<const uint32_t*>(p);
    const int magic = (int)be32_word(pw), natoms = (int)be32_word(pw + 1);
    step[f] = (int)be32_word(pw + 2);
    time[f] = __uint_as_float(be32_word(pw + 3));
    for (int k = 0; k < 9; ++k) box[9 * f + k] = __uint_as_float(be32_word(pw + 4 + k));
    const int lsize = (int)be32_word(pw + 13);
    if (magic != kXtcMagic) st = HR_XTC_BAD_MAGIC;
    else if (natoms != n_atoms || lsize != n_atoms) st = HR_XTC_ATOMS;
    if (st) {
        status[f] = st;
        return;
    }
    const long long avail_words = (end - beg) / 4;  // words of this frame
    if (n_atoms <= 9) {  // stored uncompressed
        prec[f] = 0.0f;
        if (14 + 3 * n_atoms > avail_words) {
            status[f] = HR_XTC_TRUNCATED;
            return;
        }
        for (int k = 0; k < 3 * n_atoms; ++k) out[k] = __uint_as_float(be32_word(pw + 14 + k));
        status[f] = HR_XTC_OK;
        return;
    }
    if (avail_words < 23) {
        status[f] = HR_XTC_TRUNCATED;
        return;
    }
    const float precision = __uint_as_float(be32_word(pw + 14));
    prec[f] = precision;
    int32_t minint[3], maxint[3];
    for (int k = 0; k < 3; ++k) {
        minint[k] = (int32_t)be32_word(pw + 15 + k);
        maxint[k] = (int32_t)be32_word(pw + 18 + k);
    }
    int smallidx = (int)be32_word(pw + 21);
    const long long nbytes = (int)be32_word(pw + 22);
    if (nbytes < 0 || 23 + (nbytes + 3) / 4 > avail_words) {
        status[f] = HR_XTC_TRUNCATED;
        return;
    }
    if (smallidx < kFirstIdx || smallidx >= kLastIdx) {
        status[f] = HR_XTC_STREAM;
        return;
    }
    uint32_t sizeint[3];
    for (int k = 0; k < 3; ++k) sizeint[k] = (uint32_t)(maxint[k] - minint[k]) + 1u;
    if (!sizeint[0] || !sizeint[1] || !sizeint[2]) {  // maxint below minint
        status[f] = HR_XTC_STREAM;
        return;
    }
    const bool large = (sizeint[0] | sizeint[1] | sizeint[2]) > 0xFFFFFFu;
    int bitsizeint[3] = {0, 0, 0};
    int bitsize = 0;
    if (large) {
        for (int k = 0; k < 3; ++k) {
            const int b = bit_length64(sizeint[k]);
            bitsizeint[k] = b > 32 ? 32 : b;
        }
    } else {
        // bit length of the product: below 2^72, as a 64-bit product and its
        // top byte when it does not fit
        const uint64_t xy = (uint64_t)sizeint[0] * sizeint[1];  // < 2^48
        const uint64_t lo = (xy & 0xFFFFFFull) * sizeint[2];     // < 2^48
        const uint64_t hi = (xy >> 24) * sizeint[2] + (lo >> 24);  // product >> 24
        bitsize = hi ? 24 + bit_length64(hi) : bit_length64(lo);
    }
    int smaller = kMagicInts[smallidx - 1 > kFirstIdx ? smallidx - 1 : kFirstIdx] / 2;
    int smallnum = kMagicInts[smallidx] / 2;
    uint32_t sizesmall = (uint32_t)kMagicInts[smallidx];
    const float inv = (float)(1.0 / (double)precision);
    const double inv_size2 = 1.0 / (double)sizeint[2], inv_size1 = 1.0 / (double)sizeint[1];
    BitReader br;
    br.start(pw + 23, (int)((nbytes + 3) / 4));
    const uint64_t first = (reinterpret_cast<uintptr_t>(out) >> 2);
    LineWriter wr;
    wr.start(s_ring + threadIdx.x * kRingStride, first, 3ull * (uint64_t)n_atoms);
    int n_out = 0;
    int run = 0;  // a clear flag bit keeps the previous run length
    while (n_out < n_atoms) {
        int32_t big[3];
        if (large) {
            for (int k = 0; k < 3; ++k) big[k] = (int32_t)br.take(bitsizeint[k]);
        } else {
            br.ints3(bitsize, sizeint[0], sizeint[1], sizeint[2], inv_size2, inv_size1, big);
        }
        for (int k = 0; k < 3; ++k) big[k] = (int32_t)((uint32_t)big[k] + (uint32_t)minint[k]);
        int is_smaller = 0;
        if (br.take(1)) {
            run = (int)br.take(5);
            is_smaller = run % 3;
            run -= is_smaller;
            is_smaller -= 1;
        }
        if (br.bad) {
            st = HR_XTC_STREAM;
            break;
        }
        if (n_out + 1 + run / 3 > n_atoms) {
            st = HR_XTC_STREAM;
            break;
        }
        if (run == 0) {
            wr.put_atom(big, inv);
            ++n_out;
        } else {
            // deltas from the atom before; the first small atom precedes the
            // large one in the output (stored swapped, for water)
            int32_t prev[3] = {big[0], big[1], big[2]};
            for (int k = 0; k < run; k += 3) {
                int32_t d[3];
                br.ints3(smallidx, sizesmall, sizesmall, sizesmall, 0.0, 0.0, d);
                for (int c = 0; c < 3; ++c)
                    prev[c] = (int32_t)((uint32_t)d[c] + (uint32_t)prev[c] - (uint32_t)smallnum);
                wr.put_atom(prev, inv);
                ++n_out;
                if (k == 0) {
                    wr.put_atom(big, inv);
                    ++n_out;
                }
            }
            if (br.bad) {
                st = HR_XTC_STREAM;
                break;
            }
        }
        wr.flush();
        smallidx += is_smaller;
        if (smallidx < kFirstIdx || smallidx >= kLastIdx) {
            st = HR_XTC_STREAM;
            break;
        }
        if (is_smaller < 0) {
            smallnum = smaller;
            smaller = smallidx > kFirstIdx ? kMagicInts[smallidx - 1] / 2 : 0;
        } else if (is_smaller > 0) {
            smaller = smallnum;
            smallnum = kMagicInts[smallidx] / 2;
        }
        sizesmall = (uint32_t)kMagicInts[smallidx];
    }
    wr.finish();
    status[f] = st;
}

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

int hr_xtc_scan(const void* data, int64_t n_bytes, int64_t* offsets, int64_t max_frames,
                int64_t* n_frames, int32_t* n_atoms) {
    clear_error();
    HR_REQUIRE(n_frames && n_atoms, "NULL pointer argument");
    HR_REQUIRE(n_bytes >= 0 && (data || n_bytes == 0), "invalid buffer");
    HR_REQUIRE(max_frames >= 0 && (offsets || max_frames == 0), "invalid offsets buffer");
    const unsigned char* p = static_cast<const unsigned char*>(data);
    long long pos = 0, count = 0;
    int natoms = -1;
    while (pos + 4 <= n_bytes) {
        if (pos + 56 > n_bytes) {
            set_error("truncated frame header at byte %lld", pos);
            return HR_ERR_INVALID;
        }
        const int magic = (int)be32(p + pos);
        const int na = (int)be32(p + pos + 4);
        if (magic != kXtcMagic) {
            set_error("not an XTC frame at byte %lld (magic %d)", pos, magic);
            return HR_ERR_INVALID;
        }
        if (na <= 0 || (natoms >= 0 && na != natoms)) {
            set_error("frame at byte %lld holds %d atoms (first frame: %d)", pos, na, natoms);
            return HR_ERR_INVALID;
        }
        natoms = na;
        long long next;
        if (na <= 9) {
            next = pos + 56 + 12LL * na;
        } else {
            if (pos + 92 > n_bytes) {
                set_error("truncated frame at byte %lld", pos);
                return HR_ERR_INVALID;
            }
            const long long nb = (int)be32(p + pos + 88);
            if (nb < 0) {
                set_error("negative byte count in frame at byte %lld", pos);
                return HR_ERR_INVALID;
            }
            next = pos + 92 + (nb + 3) / 4 * 4;
        }
        if (next > n_bytes) {
            set_error("truncated frame at byte %lld", pos);
            return HR_ERR_INVALID;
        }
        if (count < max_frames) offsets[count] = pos;
        ++count;
        pos = next;
    }
    *n_frames = count;
    *n_atoms = natoms < 0 ? 0 : natoms;
    return HR_OK;
}

int hr_xtc_decode(const void* data, int64_t n_bytes, const int64_t* offsets, int64_t n_frames,
                  int32_t n_atoms, float* xyz, float* time, int32_t* step, float* box,
                  float* prec, int32_t* status, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0 && n_bytes >= 0,
               "invalid sizes n_frames=%lld n_atoms=%d", (long long)n_frames, n_atoms);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(data && offsets && xyz && time && step && box && prec && status,
               "NULL pointer argument");
    HR_REQUIRE((reinterpret_cast<uintptr_t>(data) & 3) == 0, "data must be 4-byte aligned");
    HR_REQUIRE((reinterpret_cast<uintptr_t>(xyz) & 3) == 0, "xyz must be 4-byte aligned");
    const long long blocks = (n_frames + kWave - 1) / kWave;
    hipLaunchKernelGGL(xtc_decode_kernel, dim3((unsigned)blocks), dim3(kWave), 0,
                       (hipStream_t)stream, static_cast<const unsigned char*>(data),
                       (long long)n_bytes, reinterpret_cast<const long long*>(offsets),
                       (long long)n_frames, n_atoms, xyz, time, step, box, prec, status);
    return launch_status("xtc_decode_kernel");
}

}  // extern "C"
