// gpuinflate.hip - one raw Deflate stream decoded in parallel on an MI355X (gfx950), behind the
// reference's C interface (include/gpuinflate.h; reference gpuinflate/gpuinflate.h v1).
//
// Per segment (the planning is plan.h's; this file holds the kernels and the HIP backend):
//   1. k_find tests every bit position of the segment for a dynamic block header with
//      bit-sliced field checks over 32 positions at a time, then the code-length Kraft sum;
//      survivors are compacted per wavefront (one atomic per wave). k_verify applies zlib's
//      full header checks. Every verified header gets a Huffman table.
//   2. k_tables builds each table once, a workgroup per header, into the table cache.
//   3. k_probe resynchronises inside long stretches without a header.
//   4. k_decode runs one speculative decoder per lane, each landing exactly on a later start.
//   5. k_stored copies stored blocks; k_tile_a/b/d resolve back-references by carrying 32 KiB
//      windows (no pointer jumping in device memory) and write the bytes with their CRC-32.
#define GI_BUILD
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gpuinflate.h"
#include "deflate_core.h"
#include "plan.h"
#include "batch_queue.h"
#include "gzip_format.h"

using namespace gz;

namespace {

// A HIP runtime call that failed, or a device-side limit this library hit: the entry points
// turn it into GI_ERR_GPU with what() as the message, and leave the context usable
class GpuFault : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t rc, const char* call) {
    if (rc != hipSuccess) throw GpuFault(std::string(call) + ": " + hipGetErrorString(rc));
}
// the call's own text names it in the message
#define HIP_OK(call) hip_check((call), #call)

// wall time from construction (or restart) on, in seconds
class Stopwatch {
  public:
    Stopwatch() : t0_(std::chrono::steady_clock::now()) {}
    double seconds() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count(); }
    std::chrono::steady_clock::time_point origin() const { return t0_; }

  private:
    std::chrono::steady_clock::time_point t0_;
};

constexpr uint64_t CRC_TILE = 16384;   // bytes per CRC workgroup
constexpr int OUT_BUFS = 4;
constexpr uint64_t OUT_BUF = 16ull << 20;
constexpr uint64_t IN_BUF = 16ull << 20;
constexpr uint64_t INPUT_PAD = 4096;  // zero bytes after an input in d_in: the readers' look-ahead
constexpr uint64_t SEG_OUT_CAP = 64ull << 30;
constexpr uint64_t SEG_OUT_MIN = 8ull << 20;
constexpr uint64_t BATCH_MAX_STREAMS = 16384;  // streams in one batch segment at most
static_assert(BATCH_ALIGN % CRC_TILE == 0, "a batch's streams end on CRC tile boundaries");

unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)std::max<uint64_t>(1, (n + per - 1) / per); }

// ------------------------------------------------------------------ device helpers

// the cache slot of a header's table, by binary search over the sorted keys; null: none
__device__ __forceinline__ const Table* find_table(const Table* tabs, const uint64_t* keys, const uint32_t* slots,
                                                   int nkeys, uint64_t key) {
    int lo = 0, hi = nkeys;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < nkeys && keys[lo] == key ? tabs + slots[lo] : nullptr;
}

// inclusive prefix sum over a 64-lane wavefront
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// ------------------------------------------------------------------ 1. block finder

// Every bit position in (bit_lo, bit_hi), 8192 per workgroup (a lane per 32-bit word). Phase 1
// tests the header fields (BTYPE 2, HLIT and HDIST <= 29) for a word's 32 positions at once on
// bit planes and compacts the survivors (about 19%) into LDS; phase 2 gives every lane the same
// number of survivors and tests their code-length code for completeness with a Kraft-sum
// lookup table. Hits are compacted per wavefront: one atomic per wave.
// In the same pass over the words (scand not null), every byte position in [byte_lo, byte_hi)
// whose four bytes read as a stored block's LEN and NLEN (LEN = ~NLEN) goes to scand as
// (position << 16 | LEN): one match in 65,536 positions on compressed data, so a wavefront with
// none skips the compaction.
constexpr int FIND_WORDS = 256;
constexpr unsigned FIND_BUF = 512;  // a tile holds a few hits (more go straight out)
// phase 1's survivors of a tile: at most every other position, since a survivor's BTYPE bits
// (positions p+1, p+2) read 0 then 1 and the next position's would read 1 then something; about
// 1,800 of the 8,192 on compressed data. 9 workgroups fit a CU.
constexpr unsigned FIND_SURV = FIND_WORDS * 16;
// the finder's three counters set to zero: a kernel rather than memsets, so the finder's
// sequence can be captured as a graph
__global__ __launch_bounds__(64) void k_zero3(unsigned* a, unsigned* b, unsigned* c) {
    if (threadIdx.x == 0) *a = 0, *b = 0, *c = 0;
}

__global__ __launch_bounds__(256) void k_find(const uint32_t* w, uint64_t size, uint64_t word_lo, uint64_t nwords,
                                              uint64_t bit_lo, uint64_t bit_hi, uint64_t* cand, unsigned* ncand,
                                              unsigned maxcand, uint64_t byte_lo, uint64_t byte_hi, uint64_t* scand,
                                              unsigned* nscand, unsigned maxscand) {
    __shared__ uint32_t words[FIND_WORDS + 4];
    __shared__ uint16_t surv[FIND_SURV];
    __shared__ uint8_t k4[4096];  // kraft4, saturated at 255
    __shared__ uint64_t lh[FIND_BUF];  // hits of this workgroup not yet written out
    __shared__ unsigned s_n, s_lh, s_base;
    const uint32_t tid = threadIdx.x;
    Bits in;
    in.setup(w, size);
    for (uint32_t k = tid; k < 4096; k += 256) k4[k] = (uint8_t)min(kraft4(k), 255u);
    if (tid == 0) s_lh = 0;
    // the stored-block search reaches a few bytes past the bit range: its words from word_lo
    const uint64_t nwords_s = scand && byte_hi > 4 * word_lo ? (byte_hi - 4 * word_lo + 3) / 4 : 0;
    const uint64_t ntiles = ((nwords > nwords_s ? nwords : nwords_s) + FIND_WORDS - 1) / FIND_WORDS;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t w0 = word_lo + tile * FIND_WORDS;
        __syncthreads();  // the previous tile's LDS is no longer read
        for (uint32_t k = tid; k < FIND_WORDS + 4; k += 256) words[k] = in.word(w0 + k);
        if (tid == 0) s_n = 0;
        __syncthreads();
        // phase 1
        uint32_t m = 0;
        const uint64_t word = w0 + tid;
        if (tile * FIND_WORDS + tid < nwords) {
            uint64_t lo = (uint64_t)words[tid] | ((uint64_t)words[tid + 1] << 32);
#define GZ_PLANE(k) ((uint32_t)(lo >> (k)))
            m = ~GZ_PLANE(1) & GZ_PLANE(2);
            m &= ~(GZ_PLANE(4) & GZ_PLANE(5) & GZ_PLANE(6) & GZ_PLANE(7));
            m &= ~(GZ_PLANE(9) & GZ_PLANE(10) & GZ_PLANE(11) & GZ_PLANE(12));
#undef GZ_PLANE
            uint64_t base = word * 32;
            if (base <= bit_lo) m &= bit_lo - base >= 31 ? 0u : ~0u << (uint32_t)(bit_lo - base + 1);
            if (base + 32 > bit_hi) m &= bit_hi <= base ? 0u : (uint32_t)((1ull << (bit_hi - base)) - 1);
        }
        uint32_t cnt = __popc(m);
        uint32_t incl = wave_inclusive_sum(cnt);
        uint32_t wtotal = __shfl(incl, 63, 64);
        unsigned wbase = 0;
        if ((tid & 63) == 63 && wtotal) wbase = atomicAdd(&s_n, wtotal);
        wbase = __shfl(wbase, 63, 64);
        unsigned at = wbase + incl - cnt;
        while (m) {
            int o = __builtin_ctz(m);
            m &= m - 1;
            surv[at++] = (uint16_t)(tid * 32 + o);
        }
        if (scand) {
            // stored blocks: LEN/NLEN at the word's four byte positions
            uint32_t sh = 0, slen[4] = {0, 0, 0, 0};
            if (tile * FIND_WORDS + tid < nwords_s) {
                const uint64_t v = (uint64_t)words[tid] | ((uint64_t)words[tid + 1] << 32);
                for (int o = 0; o < 4; o++) {
                    const uint64_t p = word * 4 + (uint64_t)o;
                    const uint32_t len = (uint32_t)(v >> (8 * o)) & 0xFFFF, nlen = (uint32_t)(v >> (8 * o + 16)) & 0xFFFF;
                    if (p >= byte_lo && p < byte_hi && p + 4 <= size && len == (~nlen & 0xFFFFu)) {
                        sh |= 1u << o;
                        slen[o] = len;
                    }
                }
            }
            if (__ballot(sh != 0)) {
                const uint32_t sc = __popc(sh);
                const uint32_t si = wave_inclusive_sum(sc);
                const uint32_t stot = __shfl(si, 63, 64);
                unsigned sb = 0;
                if ((tid & 63) == 63) sb = atomicAdd(nscand, stot);
                sb = __shfl(sb, 63, 64);
                unsigned q = sb + si - sc;
                for (int o = 0; o < 4; o++) {
                    if (!(sh & (1u << o))) continue;
                    if (q < maxscand) scand[q] = ((word * 4 + (uint64_t)o) << 16) | slen[o];
                    q++;
                }
            }
        }
        __syncthreads();
        // phase 2
        const unsigned n = s_n;
        const uint64_t bit0 = w0 * 32;
        for (unsigned base = 0; base < n; base += 256) {
            unsigned i = base + tid;
            bool hit = false;
            uint32_t off = 0;
            if (i < n) {
                // the 64 bits from HCLEN (bit 13) and the code-length code's completeness
                off = surv[i];
                const uint32_t sb = off + 13, wi = sb >> 5, o = sb & 31;
                const uint64_t lo = (uint64_t)words[wi] | ((uint64_t)words[wi + 1] << 32);
                const uint64_t g = o ? (lo >> o) | ((uint64_t)words[wi + 2] << (64 - o)) : lo;
                hit = cl_code_complete_g(g, k4);
            }
            uint64_t mask = __ballot(hit);
            if (!mask) continue;
            const int leader = __ffsll((long long)mask) - 1;
            unsigned lbase = 0;
            if ((int)(tid & 63) == leader) lbase = atomicAdd(&s_lh, (unsigned)__popcll(mask));
            lbase = __shfl(lbase, leader, 64);
            if (hit) {
                unsigned r = lbase + __popcll(mask & ((1ull << (tid & 63)) - 1));
                if (r < FIND_BUF) {
                    lh[r] = bit0 + off;
                } else {  // buffer full (adversarial input): straight out
                    unsigned g = atomicAdd(ncand, 1u);
                    if (g < maxcand) cand[g] = bit0 + off;
                }
            }
        }
        // write the buffered hits out, one global atomic per flush, when half full or at the end
        __syncthreads();
        const unsigned nl = s_lh < FIND_BUF ? s_lh : FIND_BUF;
        if (nl && (nl >= FIND_BUF / 2 || tile + gridDim.x >= ntiles)) {
            if (tid == 0) s_base = atomicAdd(ncand, nl);
            __syncthreads();
            const unsigned gb = s_base;
            for (unsigned k = tid; k < nl; k += 256)
                if (gb + k < maxcand) cand[gb + k] = lh[k];
            __syncthreads();
            if (tid == 0) s_lh = 0;
        }
    }
}

// zlib's whole header check on the survivors: plausible_dynamic_header_lean's predicate, one
// code-length symbol a step. A wavefront takes an equal share of the candidates, and its lanes
// the next of them as they finish, so its time follows its lanes' mean parse rather than its
// longest (random candidates reject after 44 symbols on average, but the longest of 64 takes
// about 200). A lane that needs a candidate, or has used its staged input words, waits for the
// next refill: when VQ_REFILL lanes wait, or none runs, the waiting lanes stage their words and
// set up at once. A lane's next candidate is loaded one refill ahead.
constexpr int VQ_SW = 24;      // input words staged per lane (LDS)
// the header field (in order) that holds each code-length symbol's length
constexpr uint8_t CL_FIELD[19] = {3, 17, 15, 13, 11, 9, 7, 5, 4, 6, 8, 10, 12, 14, 16, 18, 0, 1, 2};
constexpr bool cl_field_ok() {
    constexpr uint8_t ord[19] = GZ_ORDER;
    for (int s = 0; s < 19; s++)
        if (ord[CL_FIELD[s]] != s) return false;
    return true;
}
static_assert(cl_field_ok(), "CL_FIELD inverts the code-length order");
constexpr int VQ_REFILL = 24;  // lanes waiting for a refill before one is made
__global__ __launch_bounds__(256) void k_verify(const uint32_t* w, uint64_t size, const uint64_t* cand,
                                                const unsigned* ncand_p, unsigned maxcand, uint64_t* hits,
                                                unsigned* nhits, unsigned maxhits) {
    __shared__ uint32_t stg[VQ_SW * 256];  // [word][lane]
    __shared__ uint8_t syms[20 * 256];     // [index][lane]: the code-length code's symbols in canonical order
    enum { IDLE, ACTIVE, PAUSED, DONE };
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const unsigned ncand = min(*ncand_p, maxcand);
    // this wavefront's share [vcur, vend)
    const uint64_t gw = (uint64_t)blockIdx.x * 4 + (tid >> 6), nwv = (uint64_t)gridDim.x * 4;
    unsigned vcur = (unsigned)(gw * ncand / nwv);
    const unsigned vend = (unsigned)((gw + 1) * ncand / nwv);
    const uint64_t nfull = size >> 2, nbits = size * 8;
    const uint32_t r = (uint32_t)(size & 3), tail_mask = r ? (uint32_t)((1ull << (8 * r)) - 1) : 0u;
    const bool aligned = ((uintptr_t)w & 15) == 0;
    int st = IDLE;
    uint64_t bit = 0, sbase = 0, rd = 0, buf = 0;
    // the lane's next candidate (nk < vend: valid)
    unsigned nk = vcur + lane;
    uint64_t nbit = nk < vend ? cand[nk] : 0;
    vcur += 64;
    uint32_t cnt = 0;
    // the code-length code: ends of lengths 1..6 (left-justified to 7 bits), and per length the
    // offset from a code to its index in canonical order (+ 128, a byte each)
    uint32_t lim1 = 0, lim2 = 0, lim3 = 0, lim4 = 0, lim5 = 0, lim6 = 0;
    uint64_t basepk = 0;
    int i = 0, total = 0, hlit = 0, prev = 0, eob = 0, nz0 = 0, nz1 = 0, ones0 = 0, ones1 = 0;
    uint32_t ks0 = 0, ks1 = 0;  // Kraft sums in units of 2^-15
    auto pull = [&]() {
        if (cnt < 32 && rd < sbase + VQ_SW) {
            buf |= (uint64_t)stg[(uint32_t)(rd - sbase) * 256 + tid] << cnt;
            rd++;
            cnt += 32;
        }
    };
    auto take = [&](uint32_t n) {
        if (cnt < n) pull();
        const uint32_t v = (uint32_t)(buf & ((1ull << n) - 1));
        buf >>= n;
        cnt -= n;
        return v;
    };
    for (;;) {
        const uint64_t need = __ballot(st == IDLE || st == PAUSED), act = __ballot(st == ACTIVE);
        if (!need && !act) break;
        if (need && (__popcll(need) >= VQ_REFILL || !act)) {
            // the refill: every idle lane takes its next candidate and loads the one after
            const uint64_t idl = __ballot(st == IDLE);
            if (st == IDLE) {
                if (nk < vend) {
                    bit = nbit;
                    nk = vcur + (unsigned)__popcll(idl & ((1ull << lane) - 1));
                    nbit = nk < vend ? cand[nk] : 0;
                } else {
                    st = DONE;
                }
            }
            vcur += (unsigned)__popcll(idl);
            // stage the words from the candidate's (idle) or the next one to read (paused)
            if (st == IDLE || st == PAUSED) {
                sbase = (st == IDLE ? bit >> 5 : rd) & ~3ull;
                if (aligned && sbase + VQ_SW <= nfull) {
                    const Q4* q = (const Q4*)w + (sbase >> 2);
#pragma unroll
                    for (int c = 0; c < VQ_SW / 4; c++) {
                        const Q4 v = q[c];
                        stg[(4 * c) * 256 + tid] = v.x;
                        stg[(4 * c + 1) * 256 + tid] = v.y;
                        stg[(4 * c + 2) * 256 + tid] = v.z;
                        stg[(4 * c + 3) * 256 + tid] = v.w;
                    }
                } else {
                    for (int k = 0; k < VQ_SW; k++) {
                        const uint64_t wi = sbase + k;
                        stg[k * 256 + tid] = wi < nfull ? w[wi] : wi == nfull && tail_mask ? w[wi] & tail_mask : 0u;
                    }
                }
                if (st == PAUSED) st = ACTIVE;
            }
            if (st == IDLE) {
                // the header up to its code lengths (plausible_dynamic_header_lean's first part)
                rd = bit >> 5;
                const uint32_t k0 = (uint32_t)(rd - sbase) * 256 + tid;  // two words, whatever cnt
                buf = ((uint64_t)stg[k0] | (uint64_t)stg[k0 + 256] << 32) >> (uint32_t)(bit & 31);
                cnt = 64 - (uint32_t)(bit & 31);
                rd += 2;
                bool ok = header_fields_ok(buf & 0x1FFFF);
                buf >>= 3;
                cnt -= 3;
                hlit = (int)take(5) + 257;
                const int hd = (int)take(5) + 1;
                const uint32_t nf = 3 * (take(4) + 4);  // bits of 3-bit lengths, in header order
                const uint32_t n1 = nf < 24 ? nf : 24, n2 = nf - n1 < 24 ? nf - n1 : 24;
                uint64_t f = take(n1);
                f |= (uint64_t)take(n2) << 24;
                f |= (uint64_t)take(nf - n1 - n2) << 48;
                // codes per length (a byte each), completeness, and the decoding tables
                uint64_t cpk = 0;
#pragma unroll
                for (int k = 0; k < 19; k++) cpk += 1ull << (8 * ((f >> (3 * k)) & 7));
                int left = 1;
                uint32_t first = 0, lims[7];
                int offs[8];
                offs[1] = 0;
                basepk = 0;
#pragma unroll
                for (int l = 1; l < 8; l++) {
                    const int c = (int)((cpk >> (8 * l)) & 0xFF);
                    left = 2 * left - c;
                    ok = ok && left >= 0;
                    lims[l - 1] = (first + c) << (7 - l);
                    basepk |= (uint64_t)(uint8_t)(offs[l] - (int)first + 128) << (8 * l);
                    first = (first + c) << 1;
                    if (l < 7) offs[l + 1] = offs[l] + c;
                }
                ok = ok && left == 0;
                lim1 = lims[0], lim2 = lims[1], lim3 = lims[2], lim4 = lims[3], lim5 = lims[4], lim6 = lims[5];
                // symbols in canonical order: symbol order[k] has the k-th length field
                uint64_t opk = 0;
#pragma unroll
                for (int l = 1; l < 8; l++) opk |= (uint64_t)offs[l] << (8 * l);
#pragma unroll
                for (int sym = 0; sym < 19; sym++) {
                    const uint32_t l = (uint32_t)((f >> (3 * CL_FIELD[sym])) & 7);
                    if (l) {
                        const uint32_t j = (uint32_t)((opk >> (8 * l)) & 0xFF);
                        syms[j * 256 + tid] = (uint8_t)sym;
                        opk += 1ull << (8 * l);
                    }
                }
                total = hlit + hd;
                i = 0, prev = 0, eob = 0, nz0 = 0, nz1 = 0, ones0 = 0, ones1 = 0, ks0 = 0, ks1 = 0;
                st = ok ? ACTIVE : IDLE;
            }
        }
        if (st == ACTIVE) {
            pull();
            if (cnt < 14) {
                st = PAUSED;  // its staged words are used up
            } else {
                // one code-length symbol (the code is complete: every 7-bit pattern starts a code)
                const uint32_t rev = __builtin_bitreverse32((uint32_t)buf) >> 25;
                const uint32_t len = 1 + (rev >= lim1) + (rev >= lim2) + (rev >= lim3) + (rev >= lim4) + (rev >= lim5) +
                                     (rev >= lim6);
                const int idx = (int)((basepk >> (8 * len)) & 0xFF) - 128 + (int)(rev >> (7 - len));
                const int s = syms[idx * 256 + tid];
                buf >>= len;
                cnt -= len;
                const uint32_t xb = s < 16 ? 0u : s == 16 ? 2u : s == 17 ? 3u : 7u;
                const int xv = (int)(buf & ((1u << xb) - 1));
                buf >>= xb;
                cnt -= xb;
                const int rep = s < 16 ? 1 : (s == 18 ? 11 : 3) + xv;
                const int val = s < 16 ? s : s == 16 ? prev : 0;
                bool bad = (s == 16 && i == 0) || i + rep > total;
                if (val) {
                    const int n0 = min(max(hlit - i, 0), rep), n1 = rep - n0;
                    ks0 += (uint32_t)n0 << (15 - val);
                    ks1 += (uint32_t)n1 << (15 - val);
                    nz0 += n0;
                    nz1 += n1;
                    if (val == 1) {
                        ones0 += n0;
                        ones1 += n1;
                    }
                    bad = bad || ks0 > 32768u || ks1 > 32768u;
                }
                if (i <= 256 && 256 < i + rep) eob = val;
                i += rep;
                prev = val;
                if (bad) {
                    st = IDLE;
                } else if (i == total) {
                    const bool acc = eob != 0 && (ks0 == 32768u || (nz0 == 1 && ones0 == 1)) &&
                                     (ks1 == 32768u || nz1 == 0 || (nz1 == 1 && ones1 == 1)) && rd * 32 - cnt <= nbits;
                    if (acc) {
                        const unsigned k = atomicAdd(nhits, 1u);
                        if (k < maxhits) hits[k] = bit;
                    }
                    st = IDLE;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ 2. tables

// One workgroup per header: lane 0 parses the code lengths, the workgroup fills the lookup
// levels in LDS and writes them out coalesced. key UINT64_MAX: the fixed code.
// keys[i] goes to slot slots[i] (slots null: base + i); n keys, or *n_p (at most nmax) when given.
// Workgroups stride over the keys.
// ests (optional): ests[i] = the key's estimated output bytes per input byte, x256 (see step 2)
__global__ __launch_bounds__(64) void k_tables(const uint32_t* w, uint64_t size, const uint64_t* keys,
                                               const uint32_t* slots, int n, const unsigned* n_p, unsigned nmax,
                                               uint32_t base, Table* tabs, uint32_t* ests) {
    static constexpr uint16_t lbase[29] = GZ_LBASE;
    static constexpr uint8_t lext[29] = GZ_LEXT;
    static constexpr uint8_t dext[30] = GZ_DEXT;
    __shared__ uint8_t lens[320];
    __shared__ uint16_t codes[320];
    __shared__ uint16_t lit[1 << LIT_BITS];
    __shared__ uint32_t dist[1 << DIST_BITS];
    __shared__ uint8_t work[64];             // the code-length code (the lean header parse's)
    __shared__ uint32_t cnt[2][16];          // codes per length: literal/length, distance
    __shared__ uint32_t offs[2][16], first[2][16];
    __shared__ int s_status, s_hlit, s_hdist;
    __shared__ uint32_t s_hdr_bits;
    __shared__ unsigned long long s_acc[5];  // the ratio estimate's sums (step 2)
    // the stream's words from the header's on, read by the whole workgroup at once (the most a
    // header parse can read is 17 + 19 * 3 + 316 * 14 bits, under HDR_WORDS words), so lane 0
    // parses without waiting on device memory
    constexpr int HDR_WORDS = 192;
    __shared__ uint32_t hw[HDR_WORDS];
    const int lane = (int)threadIdx.x;
    const int cnt_keys = n_p ? (int)min(*n_p, nmax) : n;
    for (int i = blockIdx.x; i < cnt_keys; i += gridDim.x) {
        __syncthreads();  // the workgroup is done with the last key's LDS
        Table* T = tabs + (slots ? slots[i] : base + (uint32_t)i);
        const uint64_t key = keys[i];
        if (key != UINT64_MAX) {
            Bits in;
            in.setup(w, size);
            for (int k = lane; k < HDR_WORDS; k += 64) hw[k] = in.word((key >> 5) + (uint64_t)k);
        }
        if (lane < 32) (&cnt[0][0])[lane] = 0;
        if (lane < 5) s_acc[lane] = 0;
        __syncthreads();
        // 1. the code lengths (lane 0, from LDS): the lean header parse, writing them out
        if (lane == 0) {
            if (key == UINT64_MAX) {
                for (int k = 0; k < 288; k++) lens[k] = k < 144 ? 8 : k < 256 ? 9 : k < 280 ? 7 : 8;
                for (int k = 0; k < 30; k++) lens[288 + k] = 5;
                s_status = TAB_FIXED;
                s_hlit = 288;
                s_hdist = 30;
                s_hdr_bits = 3;
            } else {
                // hw holds the stream's words (key >> 5) .. + HDR_WORDS, zeros past its end
                Bits local;
                local.setup(hw, 4ull * HDR_WORDS);
                local.seek(key & 31);
                const uint32_t t = (local.peek(3) >> 1) & 3;
                int hl = 0, hd = 0;
                int32_t st = t == 1 ? TAB_FIXED : t == 0 ? TAB_STORED : TAB_BAD;
                if (t == 2 && plausible_dynamic_header_lean<Bits, true>(local, UINT64_MAX, key & 31, work, lens, &hl, &hd))
                    st = TAB_DYNAMIC;
                s_status = st;
                s_hlit = hl;
                s_hdist = hd;
                s_hdr_bits = (uint32_t)(local.pos() - (key & 31));
            }
        }
        __syncthreads();
        const int32_t status = s_status;
        const bool fill = status == TAB_DYNAMIC || key == UINT64_MAX;
        if (!fill) {
            // a header that is fixed, stored or bad keeps no lookup levels: its blocks use slot 0
            if (lane == 0) T->status = status;
            if (lane == 0 && ests) ests[i] = 0;
            continue;
        }
        const int hl = s_hlit, hd = s_hdist;
        // 2. codes per length (the parse has checked the Kraft rules), and the sums of an estimate
        //    of the block's output bytes per input byte: each symbol taken with the probability
        //    its code length implies (2^-length, in units of 2^-15), a length symbol as its mean
        //    match length, a distance as its mean bits (sums in halves, to stay whole)
        {
            unsigned long long out2 = 0, bits2 = 0, lenw = 0, dbits2 = 0, dw = 0;
            for (int k = lane; k < hl + hd; k += 64) {
                const uint32_t L = lens[k];
                if (!L) continue;
                atomicAdd(&cnt[k >= hl ? 1 : 0][L], 1u);
                const unsigned long long wgt = 1ull << (15 - L);
                if (k < hl) {
                    if (k < 256) {
                        out2 += 2 * wgt;
                        bits2 += 2 * L * wgt;
                    } else if (k == 256) {
                        bits2 += 2 * L * wgt;
                    } else if (k < 286) {
                        const int q = k - 257;
                        out2 += (2ull * lbase[q] + (1u << lext[q]) - 1) * wgt;
                        bits2 += 2ull * (L + lext[q]) * wgt;
                        lenw += wgt;
                    }
                } else if (k - hl < 30) {
                    dbits2 += 2ull * (L + dext[k - hl]) * wgt;
                    dw += wgt;
                }
            }
            if (ests) {
                atomicAdd(&s_acc[0], out2);
                atomicAdd(&s_acc[1], bits2);
                atomicAdd(&s_acc[2], lenw);
                atomicAdd(&s_acc[3], dbits2);
                atomicAdd(&s_acc[4], dw);
            }
        }
        __syncthreads();
        if (lane == 0 && ests) {
            const double bits = (double)s_acc[1] + (s_acc[4] ? (double)s_acc[2] * (double)s_acc[3] / (double)s_acc[4] : 0.0);
            const double r = bits > 0 ? 8.0 * (double)s_acc[0] / bits : 1.0;
            ests[i] = (uint32_t)(256.0 * (r < 1.0 ? 1.0 : r > 1032.0 ? 1032.0 : r));
        }
        // 3. per code: where each length's codes start (canonical values) and their place in
        //    canonical order; the longer-code levels (long_levels' limits)
        if (lane < 2) {
            const int c = lane;
            uint32_t code = 0, off = 0;
            first[c][0] = 0;
            offs[c][0] = 0;
            for (int L = 1; L < 16; L++) {
                code = (code + (L > 1 ? cnt[c][L - 1] : 0u)) << 1;
                first[c][L] = code;
                offs[c][L] = off;
                off += cnt[c][L];
            }
            const int root = c ? DIST_BITS : LIT_BITS;
            uint32_t* lim = c ? T->dlim : T->llim;
            for (int k = 0; k < 8; k++) {
                const int len = root + 1 + k;
                if (len > 15) {
                    lim[k] = 0;
                    continue;
                }
                uint32_t end = (first[c][len] + cnt[c][len]) << (15 - len);
                if (end > 32768u) end = 32768u;
                const int32_t b = (int32_t)offs[c][len] - (int32_t)first[c][len];
                lim[k] = end | ((uint32_t)(uint16_t)(int16_t)b << 16);
            }
        }
        __syncthreads();
        // 4. every symbol's code and its entry in canonical order: its rank among the symbols of
        //    its length before it, 64 symbols at a time by ballots
        for (int c = 0; c < 2; c++) {
            const int nsym = c ? hd : hl, o = c ? hl : 0;
            uint32_t before[16];
#pragma unroll
            for (int L = 0; L < 16; L++) before[L] = 0;
            for (int at = 0; at < nsym; at += 64) {
                const int sidx = at + lane;
                const uint32_t L = sidx < nsym ? lens[o + sidx] : 0u;
                uint32_t rank = 0;
#pragma unroll
                for (int q = 1; q < 16; q++) {
                    const uint64_t m = __ballot(L == (uint32_t)q);
                    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    rank = L == (uint32_t)q ? before[q] + below : rank;
                    before[q] += (uint32_t)__popcll(m);
                }
                if (sidx < nsym) {
                    if (L) {
                        codes[o + sidx] = (uint16_t)(first[c][L] + rank);
                        const uint32_t idx = offs[c][L] + rank;
                        if (c == 0)
                            T->lent[idx] = (uint16_t)lit_entry(sidx, (int)L);
                        else
                            T->dent[idx] = dist_entry(sidx, (int)L);
                    } else {
                        codes[o + sidx] = 0;
                    }
                }
            }
        }
        if (lane == 0) {
            T->status = status;
            T->hdr_bits = s_hdr_bits;
        }
        // 5. the first-level lookups, filled in LDS and written out coalesced
        for (int k = lane; k < (1 << LIT_BITS); k += 64) lit[k] = 0;
        for (int k = lane; k < (1 << DIST_BITS); k += 64) dist[k] = 0;
        __syncthreads();
        for (int sidx = lane; sidx < hl; sidx += 64) {
            const int len = lens[sidx];
            if (!len || len > LIT_BITS) continue;
            const uint32_t e = lit_entry(sidx, len);
            for (uint32_t k = reverse_bits(codes[sidx], len); k < (1u << LIT_BITS); k += 1u << len) lit[k] = (uint16_t)e;
        }
        for (int sidx = lane; sidx < hd; sidx += 64) {
            const int len = lens[hl + sidx];
            if (!len || len > DIST_BITS) continue;
            const uint32_t e = dist_entry(sidx, len);
            for (uint32_t k = reverse_bits(codes[hl + sidx], len); k < (1u << DIST_BITS); k += 1u << len) dist[k] = e;
        }
        __syncthreads();
        for (int k = lane; k < (1 << LIT_BITS); k += 64) T->lit[k] = lit[k];
        for (int k = lane; k < (1 << DIST_BITS); k += 64) T->dist[k] = dist[k];
    }
}

// ------------------------------------------------------------------ 3. probes

__global__ __launch_bounds__(256) void k_probe(const uint32_t* w, uint64_t size, const Table* tabs,
                                               const uint64_t* keys, const uint32_t* slots, int nkeys,
                                               const Probe* pts, int n, uint64_t lead, Start* probed) {
    __shared__ uint32_t stg[SW * 256];
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Probe p = pts[i];
    const Table* own = p.mode == IN_FIXED ? nullptr : find_table(tabs, keys, slots, nkeys, p.header);
    uint64_t from = p.point > lead ? p.point - lead : 0;
    probed[i] = probe(w, size, tabs, own, p.mode, p.header, from, p.point, stg + threadIdx.x, 256);
}

// ------------------------------------------------------------------ 4. decode

// One speculative decoder per lane; the first round takes jobs longest first through `order`
__global__ __launch_bounds__(256) void k_decode(const uint32_t* w, uint64_t size, const Table* tabs,
                                                const uint64_t* keys, const uint32_t* slots, int nkeys,
                                                const Start* jobs, const int* tfirst, const uint64_t* off,
                                                const uint32_t* cap, const Start* targets, int ntargets,
                                                const int* order, int a, int n, uint16_t* pool, Result* res,
                                                int stream_start, uint64_t stop, StoredList stored, const uint64_t* lims) {
    __shared__ uint32_t stg[SW * 256];
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    int k = order ? order[i] : a + i;
    Job jb;
    jb.words = w;
    jb.size = lims && lims[k] ? lims[k] : size;
    jb.s = jobs[k];
    jb.tabs = tabs;
    jb.keys = keys;
    jb.slots = slots;
    jb.nkeys = nkeys;
    uint64_t key = jb.s.mode == IN_DYNAMIC ? jb.s.header : jb.s.mode == AT_BLOCK ? jb.s.bit : UINT64_MAX;
    jb.own = key == UINT64_MAX ? nullptr : find_table(tabs, keys, slots, nkeys, key);
    int tf = tfirst[k];
    jb.targets = target_list(targets + tf, ntargets - tf);
    jb.ntargets = ntargets - tf;
    jb.out = pool + off[k];
    jb.out_index = off[k];
    jb.cap = cap[k];
    jb.stream_start = (stream_start && k == 0) || jb.s.pad[1] == 1;
    jb.stop = stop;
    jb.fixed = tabs;
    jb.hard_end = UINT64_MAX;
    jb.stg = stg + threadIdx.x;
    jb.stride = 256;
    res[k] = run(jb, &stored);
}

// The rescue's header verification (rescue.cpp:549-554, verify_at): from each candidate block
// start, a decode that writes nothing, up to VERIFY_OUT symbols; a header is confirmed when the
// decode ends the stream cleanly, or fills VERIFY_OUT (within one symbol) or crosses
// VERIFY_BLOCKS block starts. Dynamic headers it crosses come from the table cache, which holds
// every plausible header from the candidates' range to the input's end (those are all the
// dynamic headers a decode can accept: a header past them reads past the input).
constexpr uint32_t VERIFY_OUT = 256u << 10;
constexpr uint32_t VERIFY_BLOCKS = 2;

// (64-lane workgroups: a few thousand long decodes spread over every CU)
__global__ __launch_bounds__(64) void k_rescue_verify(const uint32_t* w, uint64_t size, const Table* tabs,
                                                      const uint64_t* keys, const uint32_t* slots, int nkeys,
                                                      const uint64_t* cand, int n, uint32_t* ok) {
    __shared__ uint32_t stg[SW * 64];
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    const uint64_t bit = cand[i];
    Job jb;
    jb.words = w;
    jb.size = size;
    jb.s.bit = bit;
    jb.s.header = bit;
    jb.s.mode = AT_BLOCK;
    jb.s.final = -1;
    for (int k = 0; k < 6; k++) jb.s.pad[k] = 0;
    jb.tabs = tabs;
    jb.keys = keys;
    jb.slots = slots;
    jb.nkeys = nkeys;
    jb.own = find_table(tabs, keys, slots, nkeys, bit);
    jb.targets = TargetList{};
    jb.ntargets = 0;
    jb.out = nullptr;
    jb.out_index = 0;
    jb.cap = VERIFY_OUT;
    jb.stream_start = false;
    jb.stop = UINT64_MAX;
    jb.fixed = tabs;
    jb.hard_end = UINT64_MAX;
    jb.stg = stg + threadIdx.x;
    jb.stride = 64;
    const Result r = run<false>(jb, nullptr);
    ok[i] = r.status == OK || (r.status == ERR_OVERFLOW &&
                               (r.blocks >= VERIFY_BLOCKS || r.out_len >= VERIFY_OUT - MAX_SYMBOL_OUT));
}

// Groups: L lanes decode one interval together (deflate_core.h, "groups"): the
// interval's code goes into LDS once, lanes 1.. find their starts by probing, every lane decodes
// landing on a later lane's start or on a target after the interval, and lane 0 walks the chain.
//
// The groups of a decode round are a work queue in device memory: their descriptors in the order
// the host queued them (longest interval first) and a head counter. A batch segment's streams all
// share one queue. Two launch forms:
//   LOOP false: a workgroup per GPW groups, the queue's items in workgroup order (the hardware's
//               dispatcher hands them out as compute units free up; the default);
//   LOOP true:  one persistent launch of a chosen number of workgroups per compute unit, each
//               taking the next item from the queue's head counter (advanced atomically) until
//               the queue is empty. It holds its compute units for the whole round, which the
//               next segment's look-ahead finder, running beside it, then waits for: measured
//               about 5% slower on one 4 GiB stream, so it is a schedule field the autotuner tries.
// L lanes a group (16, 32 or 64), GPW = 64 / L groups a 64-lane workgroup.
// a lane's probe gives up sooner than a stand-alone one: a lane without a start only makes the
// lane before it decode further, while a wavefront waits for its slowest probe
constexpr size_t FOLD_LANES = 4096;   // a later round's lane jobs go into the group launch up to this many
constexpr int GROUP_PROBE_TRIES = 4;  // measured: 4 beats 2, 6, 8 and 12 on level-6 data

template <int L, bool LOOP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_groups(const uint32_t* w, uint64_t size, const Table* tabs, const uint64_t* keys,
                                               const uint32_t* slots, int nkeys, const GroupDesc* groups, int ngroups,
                                               unsigned* queue_head, const Start* targets, int ntargets, uint16_t* pool,
                                               GroupResult* gres, GroupHead* ghead,
                                               int stream_start, uint64_t stop, uint64_t lead, StoredList stored,
                                               uint64_t* trace) {
    constexpr int GPW = 64 / L;
    static_assert(GPW * L == 64, "groups fill a wavefront");
    __shared__ Table tab[GPW];
    // the lanes' staging words while they decode; then their results
    __shared__ __align__(16) uint32_t stg[SW * 64];
    static_assert(sizeof(Result) * 64 <= sizeof(uint32_t) * SW * 64, "results fit in the staging words");
    Result(*lres)[L] = reinterpret_cast<Result(*)[L]>(stg);
    __shared__ Start lst[GPW][L];
    __shared__ int lane_of[GPW][L];
    __shared__ int first_t[GPW][L];
    __shared__ int s_nv[GPW], s_nt[GPW];  // valid later lane starts (lane_of), and with the targets after
    __shared__ int s_kind[GPW];         // the first block: 0 none, 1 fixed, 2 dynamic in LDS
    __shared__ uint64_t s_first[GPW];   // its first symbol
    __shared__ const Table* s_src[GPW];
    __shared__ int s_item;
    const int sub = threadIdx.x / L, lane = threadIdx.x % L;
    for (;;) {
        int item = (int)blockIdx.x;
        if (LOOP) {
            __syncthreads();  // the item before is done with the shared arrays
            if (threadIdx.x == 0) s_item = (int)atomicAdd(queue_head, 1u);
            __syncthreads();
            // the same for the whole workgroup (one wavefront): a uniform loop exit
            item = __builtin_amdgcn_readfirstlane(s_item);
            if ((int64_t)item * GPW >= ngroups) break;  // the queue is empty: every workgroup gets here
        }
        const uint64_t t0 = trace ? wall_clock64() : 0;
        const int gi = item * GPW + sub;
        const bool active = gi < ngroups;
        GroupDesc gd;
        if (active) gd = groups[gi];
        const uint64_t lim = active && gd.lim ? gd.lim : size;  // its stream's bytes
        // phase 0: the interval's first block and its code
        if (active && lane == 0) {
            Bits in;
            in.setup(w, lim);
            int kind = 0;
            uint64_t first = gd.s.bit;
            const Table* src = nullptr;
            if (gd.s.mode == IN_FIXED) {
                kind = 1;
            } else if (gd.s.mode == IN_DYNAMIC) {
                src = find_table(tabs, keys, slots, nkeys, gd.s.header);
                kind = src && src->status == TAB_DYNAMIC ? 2 : 0;
            } else {
                in.seek(gd.s.bit);
                uint32_t type = in.peek(3) >> 1;
                if (type == 1) {
                    kind = 1;
                    first = gd.s.bit + 3;
                } else if (type == 2) {
                    src = find_table(tabs, keys, slots, nkeys, gd.s.bit);
                    if (src && src->status == TAB_DYNAMIC) {
                        kind = 2;
                        first = gd.s.bit + src->hdr_bits;
                    }
                }
            }
            s_kind[sub] = kind;
            s_first[sub] = first;
            s_src[sub] = kind == 1 ? tabs : src;
        }
        __syncthreads();
        if (active && s_kind[sub]) {
            const uint4* from = (const uint4*)s_src[sub];
            uint4* to = (uint4*)&tab[sub];
            for (int k = lane; k < (int)(sizeof(Table) / 16); k += L) to[k] = from[k];
        }
        __syncthreads();
        // phase 1: lane starts
        const int kind = active ? s_kind[sub] : 0;
        const uint64_t first = active ? s_first[sub] : 0;
        const uint64_t header0 = active ? (gd.s.mode == AT_BLOCK ? gd.s.bit : gd.s.header) : 0;
        if (active) {
            Start st;
            if (lane == 0) {
                st = gd.s;
            } else {
                st.bit = UINT64_MAX;
                st.header = UINT64_MAX;
                st.mode = AT_BLOCK;
                st.final = -1;
                for (int i = 0; i < 6; i++) st.pad[i] = 0;
                if (kind && gd.split_end > first) {
                    uint64_t point = group_point(first, gd.split_end, lane, L);
                    uint64_t from = point > first + lead ? point - lead : first;
                    st = probe(w, lim, kind == 1 ? &tab[sub] : tabs, kind == 2 ? &tab[sub] : nullptr, gd.s.mode, header0,
                               from, point, stg + threadIdx.x, 64, GROUP_PROBE_TRIES);
                    // BFINAL stays unknown: a probe may have synchronised in a later block with the same
                    // code (any fixed block); the chain walk learns it from the lane before
                }
            }
            lst[sub][lane] = st;
        }
        __syncthreads();
        const uint64_t t1 = trace ? wall_clock64() : 0;
        if (active && lane == 0) {
            int nv = 0;
            uint64_t prev = gd.s.bit;
            first_t[sub][0] = 0;
            for (int i = 1; i < L; i++) {
                const Start& st = lst[sub][i];
                if (st.bit != UINT64_MAX && st.bit > prev && st.bit < gd.end) {
                    lane_of[sub][nv] = i;
                    nv++;
                    first_t[sub][i] = nv;
                    prev = st.bit;
                } else {
                    first_t[sub][i] = -1;
                }
            }
            // (a lane job run as a group, lane_cap 0, may land on any target after it)
            int ngt = ntargets - gd.tfirst;
            ngt = ngt < 0 ? 0 : gd.lane_cap && ngt > GROUP_TARGETS ? GROUP_TARGETS : ngt;
            s_nv[sub] = nv;
            s_nt[sub] = nv + ngt;
        }
        __syncthreads();
        // phase 2: every lane with a start decodes
        Result mine;
        const bool runs = active && first_t[sub][lane] >= 0;
        if (runs) {
            const Start st = lst[sub][lane];
            const int ft = first_t[sub][lane];
            Job jb;
            jb.words = w;
            jb.size = lim;
            jb.s = st;
            jb.tabs = tabs;
            jb.keys = keys;
            jb.slots = slots;
            jb.nkeys = nkeys;
            const uint64_t key = st.mode == IN_DYNAMIC ? st.header : st.mode == AT_BLOCK ? st.bit : UINT64_MAX;
            jb.own = kind == 2 && key == header0 ? &tab[sub]
                     : key == UINT64_MAX     ? nullptr
                                             : find_table(tabs, keys, slots, nkeys, key);
            // its landing list: the valid lane starts after its own (from lst, through lane_of), then the
            // targets after the interval (read where they are: lanes reach them only at block ends)
            jb.targets.a = lst[sub];
            jb.targets.ai = lane_of[sub] + ft;
            jb.targets.na = s_nv[sub] - ft;
            jb.targets.b = targets + gd.tfirst;
            jb.targets.nb = s_nt[sub] - s_nv[sub];
            jb.ntargets = s_nt[sub] - ft;
            jb.out_index = gd.pool + (uint64_t)lane * gd.lane_cap;
            jb.out = pool + jb.out_index;
            // with no split region (stored data at the interval's start) lane 0 runs alone and has
            // the whole group's room
            jb.cap = !(kind && gd.split_end > first) && lane == 0
                         ? gd.lane_cap * L + gd.last_extra
                         : gd.lane_cap + (lane == L - 1 ? gd.last_extra : 0);
            jb.stream_start = lane == 0 && ((stream_start && gd.job == 0) || gd.s.pad[1] == 1);
            jb.stop = stop;
            jb.fixed = kind == 1 ? &tab[sub] : tabs;
            jb.hard_end = gd.hard ? gd.hard : UINT64_MAX;
            jb.stg = stg + threadIdx.x;
            jb.stride = 64;
            mine = run(jb, &stored);
        }
        __syncthreads();  // staging is free
        if (runs) lres[sub][lane] = mine;
        __syncthreads();
        // phase 3: the chain
        if (active && lane == 0) {
            GroupResult g;
            group_walk(lst[sub], lres[sub], lane_of[sub], first_t[sub], s_nv[sub], L, g);
            gres[gd.job] = g;
            ghead[gd.job] = GroupHead{g.r, g.chain, g.eob_lane, g.eob_len};
            if (trace) {
                // timing of this group (10 ns ticks) and its longest lane: out_len, lane, status
                uint32_t mx = 0, ml = 0;
                for (int l = 0; l < L; l++)
                    if (first_t[sub][l] >= 0 && lres[sub][l].out_len >= mx) {
                        mx = lres[sub][l].out_len;
                        ml = (uint32_t)l;
                    }
                trace[gi * 4 + 0] = t0;
                trace[gi * 4 + 1] = t1;
                trace[gi * 4 + 2] = wall_clock64();
                trace[gi * 4 + 3] = mx | (uint64_t)ml << 32 | (uint64_t)(uint8_t)lres[sub][ml].status << 40 |
                                    (uint64_t)s_nv[sub] << 48;
            }
        }
        if (trace && active) {
            // every lane: its start, where it ended, and how (status, landed, symbols)
            uint64_t* lt = trace + 4 * (uint64_t)ngroups + ((uint64_t)gi * L + lane) * 3;
            const bool ran = first_t[sub][lane] >= 0;
            lt[0] = lst[sub][lane].bit;
            lt[1] = ran ? lres[sub][lane].end_bit : 0;
            lt[2] = ran ? (uint64_t)(uint32_t)lres[sub][lane].out_len | (uint64_t)(uint8_t)lres[sub][lane].status << 32 |
                              (uint64_t)(uint16_t)lres[sub][lane].landed << 40
                        : ~0ull;
        }
        if (!LOOP) break;
    }
}

// ------------------------------------------------------------------ 5. resolve

// the deferred stored-block copies: workgroups stride over the list (its length in device
// memory, so the host need not wait for it)
__global__ void k_stored(const StoredCopy* list, const unsigned* n_p, unsigned cap, const uint8_t* in, uint16_t* pool) {
    typedef uint2 __attribute__((aligned(1))) B8;  // 8 bytes at any address
    typedef uint4 __attribute__((aligned(2))) S8;  // 8 symbols at any symbol
    const unsigned n = min(*n_p, cap);
    for (unsigned b = blockIdx.x; b < n; b += gridDim.x) {
        StoredCopy c = list[b];
        // 8 bytes a thread a step: one load, one 16-byte store of the 8 symbols; the tail bytewise
        const uint32_t whole = c.len & ~7u;
        for (uint32_t i = 8 * threadIdx.x; i < whole; i += 8 * blockDim.x) {
            const B8 v = *(const B8*)(in + c.src + i);
            const S8 q = {(v.x & 0xFF) | (v.x & 0xFF00) << 8, (v.x >> 16 & 0xFF) | (v.x >> 8 & 0xFF0000),
                          (v.y & 0xFF) | (v.y & 0xFF00) << 8, (v.y >> 16 & 0xFF) | (v.y >> 8 & 0xFF0000)};
            *(S8*)(pool + c.dst + i) = q;
        }
        for (uint32_t i = whole + threadIdx.x; i < c.len; i += blockDim.x) pool[c.dst + i] = in[c.src + i];
    }
}

// A batch's streams packed into one input: each chunk copies len bytes from its stream (any
// alignment) to dst (16-byte aligned) and zeros the bytes after them up to zlen
struct PackChunk {
    const uint8_t* src;
    uint64_t dst;
    uint32_t len;
    uint32_t zlen;
};
constexpr uint32_t PACK_CHUNK = 256u << 10;

__global__ __launch_bounds__(256) void k_pack(const PackChunk* chunks, uint8_t* in) {
    typedef uint4 __attribute__((aligned(1))) U4;
    const PackChunk c = chunks[blockIdx.x];
    uint8_t* d = in + c.dst;
    const uint32_t n16 = c.len / 16;
    for (uint32_t i = threadIdx.x; i < n16; i += 256) *(uint4*)(d + 16 * i) = *(const U4*)(c.src + 16 * i);
    for (uint32_t i = 16 * n16 + threadIdx.x; i < c.zlen; i += 256) d[i] = i < c.len ? c.src[i] : (uint8_t)0;
}

// ------------------------------------------------------------------ CRC-32 algebra

__device__ __host__ inline uint32_t gf2_mul(uint32_t a, uint32_t b) {
    // product mod the CRC-32 polynomial, reflected bit order (bit 31 holds x^0)
    uint32_t p = 0;
    for (int i = 0; i < 32; i++) {
        if (a & (0x80000000u >> i)) p ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return p;
}

// x^(8n) mod P
uint32_t x8n(uint64_t n) {
    uint32_t r = 0x80000000u;     // x^0
    uint32_t sq = 0x00800000u;    // x^8
    while (n) {
        if (n & 1) r = gf2_mul(r, sq);
        sq = gf2_mul(sq, sq);
        n >>= 1;
    }
    return r;
}

// multiplication by a constant in the CRC's field, a byte at a time (it is linear)
struct ConstMul {
    uint32_t t[4][256];
    void set(uint32_t c) {
        for (int k = 0; k < 4; k++)
            for (uint32_t b = 0; b < 256; b++) t[k][b] = gf2_mul(b << (8 * k), c);
    }
    uint32_t mul(uint32_t a) const { return t[0][a & 255] ^ t[1][(a >> 8) & 255] ^ t[2][(a >> 16) & 255] ^ t[3][a >> 24]; }
};

// ------------------------------------------------------------------ pieces
// The segment's pieces from its chain entries (one thread per entry): pool index, length and
// output index of every piece, in output order (see ChainEntry)
__global__ __launch_bounds__(256) void k_pieces(const ChainEntry* ents, int n, const GroupResult* gres, uint64_t* poff,
                                                uint32_t* plen, uint64_t* pdst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const ChainEntry e = ents[i];
    uint32_t q = e.piece0;
    if (!e.cap) {
        poff[q] = e.off;
        plen[q] = e.len;
        pdst[q] = e.dst;
        return;
    }
    const GroupResult* g = gres + e.job;
    uint64_t d = e.dst;
    for (uint64_t c = e.chain; c; c &= c - 1) {
        const uint32_t l = (uint32_t)__builtin_ctzll(c);
        const uint32_t ll = l == e.trim_lane ? e.trim_len : g->len[l];
        poff[q] = e.off + (uint64_t)l * e.cap;
        plen[q] = ll;
        pdst[q] = d;
        d += ll;
        q++;
    }
}

// ------------------------------------------------------------------ back-references, by window carry
// Every reference reaches at most 32 KiB back, so the bytes of a stretch of output depend on the
// rest of the stream only through the 32 KiB before it. The segment's output is cut into super
// tiles (super_size() positions); their entries are resolved in three passes:
//   A  (a workgroup per super tile, all at once) walks the super tile in TILE-position tiles with
//      the 32 KiB before each tile in LDS, resolving chains inside a tile by pointer doubling
//      in LDS: every entry becomes a byte, or a reference into the 32 KiB before its super tile;
//   B  finds the 32 KiB of bytes before every super tile by pointer doubling over the super
//      tiles' windows (each window's references point into the window before it): log2 of the
//      number of super tiles rounds, each over all windows at once;
//   D  (a workgroup per CRC_TILE of output) maps entries to bytes through their super tile's
//      window, writes them, and computes each tile's CRC-32.
// Entry form between A and D ("window form", 16 bits): below 256 a byte; 256 + k the byte k of
// the 32 KiB window before the super tile.
constexpr uint32_t TILE = 16384;
// A super tile is a multiple of TILE between SUPER_MIN and SUPER_MAX positions: a segment's
// output is cut into at most one per CU (`super_split`, the device's CU count: a workgroup of
// pass A fills a CU's LDS, so a second wave of workgroups would double the serial walk), and a
// short segment still keeps many workgroups busy, each walking few tiles
constexpr uint64_t SUPER_MAX = 64ull << 20;
constexpr uint64_t SUPER_MIN = 4 * TILE;
constexpr uint64_t SUPER_SPLIT = 256;  // at least this many super tiles' windows are allocated
static_assert(SUPER_MIN % TILE == 0 && SUPER_MIN >= 2 * WINDOW, "super tiles hold whole tiles and a window");
static_assert(TILE % CRC_TILE == 0, "CRC tiles lie in one super tile");
inline uint64_t super_size(uint64_t total, uint64_t super_split) {
    const uint64_t s = (total / super_split + TILE - 1) / TILE * TILE;
    return std::clamp<uint64_t>(s, SUPER_MIN, SUPER_MAX);
}
// pass A working form: INTILE | 4 j, a reference to tile entry j (the low 16 bits are its byte
// offset in the tile's LDS array); below INTILE, window form
constexpr uint32_t INTILE = 0x10000u;
constexpr uint32_t TA_THREADS = 1024;
constexpr uint32_t TA_PER = TILE / TA_THREADS;  // tile positions per thread
constexpr uint32_t TA_PIECES = TA_THREADS;      // piece descriptors a tile reads at once
// the value unchanged, but opaque to the compiler: keeps it from holding sixteen loop-invariant
// 64-bit copies of a thread's positions in registers across the tile loop
__device__ __forceinline__ uint32_t ta_opaque(uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
}

__global__ __launch_bounds__(TA_THREADS) void k_tile_a(const uint16_t* pool, const uint64_t* poff, const uint32_t* plen,
                                                const uint64_t* pdst, int np, uint64_t total, uint64_t super,
                                                uint16_t* wf, uint64_t* trace) {
    // trace (optional): per workgroup, 10 ns ticks spent reading pieces, filling the tile, in
    // chains and writing, and the doubling rounds
    uint64_t tr[4] = {0, 0, 0, 0}, rounds = 0, tk = trace ? wall_clock64() : 0;
    __shared__ uint16_t win[WINDOW];  // window form of the 32 KiB before the tile, a ring from wbase
    __shared__ uint32_t cur[TILE];
    __shared__ uint64_t s_pd[TA_PIECES];  // the tile's pieces: first position, and pool index - position
    __shared__ int64_t s_dl[TA_PIECES];
    __shared__ uint32_t s_cnt[TA_THREADS / 64];  // per wavefront: pieces over the tile, and (<< 16) to its end
    __shared__ int s_q, s_flag[3];
    const uint32_t tid = threadIdx.x;
    const uint64_t s0 = blockIdx.x * super;
    const uint64_t s1 = s0 + super < total ? s0 + super : total;
    uint32_t wbase = 0;  // logical window index k lives at win[(wbase + k) & (WINDOW - 1)]
    for (uint32_t k = tid; k < WINDOW; k += TA_THREADS) win[k] = (uint16_t)(256 + k);
    // The tiles are pipelined: a tile's pool entries are loaded (into e) while the tile before
    // resolves its chains, and the piece descriptors a tile starts from one tile earlier still
    // (q_pre: the tile's first piece, -1 when not known; d_pre, o_pre: this lane's descriptor)
    uint32_t e[TA_PER];
    int q_pre = -1;
    uint64_t d_pre = UINT64_MAX;
    int64_t o_pre = 0;
    auto load_tile = [&](uint64_t ts) {
        const uint32_t n = (uint32_t)(s1 - ts < TILE ? s1 - ts : TILE);
        int q0 = q_pre;
        uint64_t d = d_pre;
        int64_t o = o_pre;
        if (q0 < 0) {
            // the last piece that starts at or before ts, by a binary search (a super tile's
            // first tile, or after a tile of more pieces than descriptors)
            if (tid == 0) {
                int lo = 0, hi = np;
                while (lo < hi) {
                    int mid = (lo + hi) >> 1;
                    if (pdst[mid] <= ts)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                s_q = lo - 1;
            }
            __syncthreads();
            q0 = s_q;
            d = q0 + (int)tid < np ? pdst[q0 + tid] : UINT64_MAX;
            o = q0 + (int)tid < np ? (int64_t)poff[q0 + tid] : 0;
        }
        // the pieces over the tile (they follow each other without gaps), read at once, the
        // slots after them marked by a first position no tile reaches; and how many start at or
        // before the next tile, whose last is that tile's first (counted by wavefront ballots)
        const bool in = q0 + (int)tid < np;
        const bool have = in && (tid == 0 || d < ts + n);
        const bool le = in && (tid == 0 || d <= ts + n);
        s_pd[tid] = have ? d : UINT64_MAX;
        s_dl[tid] = o - (int64_t)d;
        const uint64_t bh = __ballot(have), bl = __ballot(le);
        if ((tid & 63) == 0) s_cnt[tid >> 6] = (uint32_t)__popcll(bh) | (uint32_t)__popcll(bl) << 16;
        __syncthreads();
        uint32_t cs = 0;
#pragma unroll
        for (uint32_t w = 0; w < TA_THREADS / 64; w++) cs += s_cnt[w];
        const int nb = (int)(cs & 0xFFFF), nle = (int)(cs >> 16);
        q_pre = nle < (int)TA_PIECES ? q0 + nle - 1 : -1;
        if (q_pre >= 0) {
            d_pre = q_pre + (int)tid < np ? pdst[q_pre + tid] : UINT64_MAX;
            o_pre = q_pre + (int)tid < np ? (int64_t)poff[q_pre + tid] : 0;
        }
        const bool more = nb == (int)TA_PIECES && q0 + nb < np && pdst[q0 + nb] < ts + n;
        if (!more) {
            // each thread: positions tid + m * TA_THREADS, its piece found by walking forward
            // (the next piece's first position and this piece's offset kept in registers); the
            // loads are only waited for when the entries are used
            int bq = 0;
            uint64_t nxt = s_pd[1];
            int64_t dl = s_dl[0];
#pragma unroll
            for (uint32_t m = 0; m < TA_PER; m++) {
                const uint32_t i = ta_opaque(tid + m * TA_THREADS);
                const uint64_t p = ts + i;
                while (p >= nxt) {
                    bq++;
                    dl = s_dl[bq];
                    nxt = bq + 1 < (int)TA_PIECES ? s_pd[bq + 1] : UINT64_MAX;
                }
                e[m] = i < n ? pool[(uint64_t)((int64_t)p + dl)] : 0u;
            }
        } else {
            // more pieces than descriptors: one piece after another
            for (uint32_t m = 0; m < TA_PER; m++) e[m] = 0;
            for (int q = q0; q < np && pdst[q] < ts + n; q++) {
                const uint64_t pd = pdst[q];
                const uint64_t a = pd > ts ? pd : ts;
                const uint64_t pe = pd + plen[q];
                const uint64_t bnd = pe < ts + n ? pe : ts + n;
                for (uint32_t m = 0; m < TA_PER; m++) {
                    const uint64_t p = ts + ta_opaque(tid + m * TA_THREADS);
                    if (p >= a && p < bnd) e[m] = pool[poff[q] + (p - pd)];
                }
            }
        }
    };
    if (s0 < s1) load_tile(s0);
    if (trace) {
        const uint64_t now = wall_clock64();
        tr[0] += now - tk;
        tk = now;
    }
    for (uint64_t ts = s0; ts < s1; ts += TILE) {
        const uint32_t n = (uint32_t)(s1 - ts < TILE ? s1 - ts : TILE);
        uint32_t c[TA_PER];  // cur[i] of the thread's own positions (no other lane writes them)
#pragma unroll
        for (uint32_t m = 0; m < TA_PER; m++) {
            const uint32_t i = tid + m * TA_THREADS;
            c[m] = 0;
            if (i < n) {
                const uint32_t v = e[m];
                uint32_t x;
                if (v <= REF) {
                    x = v;
                } else {
                    const uint32_t d = v - REF;
                    x = d <= i ? (INTILE | (i - d) << 2) : win[(wbase + WINDOW - (d - i)) & (WINDOW - 1)];
                }
                cur[i] = x;
                c[m] = x;
            }
        }
        if (tid == 0) s_flag[0] = s_flag[1] = s_flag[2] = 0;
        __syncthreads();
        if (trace) {
            const uint64_t now = wall_clock64();
            tr[3] += now - tk;
            tk = now;
        }
        if (ts + TILE < s1) load_tile(ts + TILE);
        if (trace) {
            const uint64_t now = wall_clock64();
            tr[0] += now - tk;
            tk = now;
        }
        // chains inside the tile: pointer doubling until every entry is in window form (each
        // reference points to a smaller index, so a chain halves each round; reading an entry
        // another lane rewrote this round only moves further along the same chain). One barrier
        // a round: round r reports in s_flag[r % 3] and clears the flag of round r + 1.
        // A thread takes its entries in order, so an entry can build on the one it just
        // resolved (references point back); an entry is pending while its own value (in c,
        // which no other lane writes) is in tile form.
        for (int r = 0;; r++) {
            uint32_t still = 0;
#pragma unroll
            for (uint32_t m = 0; m < TA_PER; m++) {
                if (c[m] >= INTILE) {
                    const uint32_t f = *(const uint32_t*)((const uint8_t*)cur + (c[m] & 0xFFFC));
                    cur[tid + m * TA_THREADS] = f;
                    c[m] = f;
                    still |= f;
                }
            }
            if (still >= INTILE) s_flag[r % 3] = 1;
            if (tid == 0) s_flag[(r + 1) % 3] = 0;
            __syncthreads();
            rounds++;
            if (!s_flag[r % 3]) break;
        }
        if (trace) {
            const uint64_t now = wall_clock64();
            tr[1] += now - tk;
            tk = now;
        }
#pragma unroll
        for (uint32_t m = 0; m < TA_PER; m++) {
            const uint32_t i = tid + m * TA_THREADS;
            if (i < n) {
                const uint16_t x = (uint16_t)c[m];
                wf[ts + ta_opaque(i)] = x;
                win[(wbase + i) & (WINDOW - 1)] = x;  // the tile's entries replace the oldest n
            }
        }
        wbase = (wbase + n) & (WINDOW - 1);
        __syncthreads();
        if (trace) {
            const uint64_t now = wall_clock64();
            tr[2] += now - tk;
            tk = now;
        }
    }
    if (trace && tid == 0) {
        trace[blockIdx.x * 5 + 0] = tr[0];
        trace[blockIdx.x * 5 + 1] = tr[1];
        trace[blockIdx.x * 5 + 2] = tr[2];
        trace[blockIdx.x * 5 + 3] = rounds;
        trace[blockIdx.x * 5 + 4] = tr[3];
    }
}

// Pass B. A window entry is a byte (WB_BYTE | b) or a reference (j << 15 | k): byte k of the
// window before super tile j. Window 0 is the segment's own (the W bytes before it, zeros
// before the stream's start); window s >= 1 starts as the last 32 KiB of super tile s-1, whose
// references point into window s-1.
constexpr uint32_t WB_BYTE = 0x80000000u;

__global__ __launch_bounds__(256) void k_win_init(const uint16_t* wf, const uint8_t* segwin, uint32_t W, int nsuper,
                                                  uint64_t super, uint32_t* wref) {
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t s = (uint32_t)(g / WINDOW), k = (uint32_t)(g % WINDOW);
    if (s >= (uint32_t)nsuper) return;
    uint32_t x;
    if (s == 0) {
        x = WB_BYTE | (k >= WINDOW - W ? segwin[k - (WINDOW - W)] : 0u);
    } else {
        const uint32_t v = wf[(uint64_t)s * super - WINDOW + k];
        if (v < 256)
            x = WB_BYTE | v;
        else if (s == 1)
            x = WB_BYTE | (v - 256 >= WINDOW - W ? segwin[v - 256 - (WINDOW - W)] : 0u);
        else
            x = (s - 1) << 15 | (v - 256);
    }
    wref[g] = x;
}

// one doubling round, in place (an entry read after another lane rewrote it this round is only
// further along the same chain)
__global__ __launch_bounds__(256) void k_win_jump(int nsuper, uint32_t* wref) {
    const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= (uint64_t)nsuper * WINDOW) return;
    const uint32_t x = wref[g];
    if (x & WB_BYTE) return;
    wref[g] = wref[(uint64_t)(x >> 15) * WINDOW + (x & (WINDOW - 1))];
}

__global__ __launch_bounds__(256) void k_win_bytes(int nsuper, const uint32_t* wref, uint8_t* wins) {
    const uint64_t g = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (g >= (uint64_t)nsuper * WINDOW) return;
    const uint4 x = *(const uint4*)(wref + g);
    *(uint32_t*)(wins + g) = (x.x & 255) | (x.y & 255) << 8 | (x.z & 255) << 16 | (x.w & 255) << 24;
}

// Pass D: bytes and their CRC, one workgroup per CRC_TILE of output. Entries are read 8 at a
// time (coalesced), a window byte only where an entry names one. Each thread then takes the CRC
// of PER consecutive bytes with 4-byte table steps (its row padded by a word in LDS, so the
// threads' reads fall in different banks); each row's CRC is shifted past the rows after it
// and all are XORed together (one multiplication a thread, a wavefront reduction). Raw CRC
// register (no inversions); a tile shorter than the rest is aligned to its end (leading zeros
// leave a raw CRC unchanged).
constexpr uint32_t TD_PER = CRC_TILE / 256;  // bytes per thread
constexpr uint32_t TD_ROW = TD_PER / 4 + 1;  // words per thread's row in LDS

// A batch (ns > 0) writes each tile to the stream that owns it: the last of the ns streams whose
// first byte sbeg[] (ascending) lies before the tile's end; its bytes go to sdst[] from sbeg on,
// and the zeros before sbeg are not written. A batch's total is a multiple of CRC_TILE.
__global__ __launch_bounds__(256) void k_tile_d(const uint16_t* wf, uint64_t total, uint64_t super, const uint8_t* wins,
                                                uint8_t* out, const uint32_t* crc4, const uint32_t* shifts,
                                                uint32_t* tile_crc, const uint64_t* sbeg, const uint64_t* sdst, int ns) {
    typedef uint2 __attribute__((aligned(1))) U2;
    __shared__ uint32_t tab[4][256];
    __shared__ uint32_t rows[256 * TD_ROW];
    __shared__ uint32_t part[256];
    __shared__ int s_str;
    const uint32_t tid = threadIdx.x;
    const uint64_t t0 = blockIdx.x * CRC_TILE;
    const uint64_t len = total - t0 < CRC_TILE ? total - t0 : CRC_TILE;
    const uint32_t pad = (uint32_t)(CRC_TILE - len);
    const uint8_t* ws = wins + (t0 / super) * WINDOW;
    if (ns && tid == 0) {
        int lo = 0, hi = ns;  // the first stream starting at or after the tile's end
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (sbeg[mid] < t0 + CRC_TILE)
                lo = mid + 1;
            else
                hi = mid;
        }
        s_str = lo - 1;
    }
    for (uint32_t k = tid; k < 1024; k += 256) (&tab[0][0])[k] = crc4[k];
    uint64_t first = 0;  // a batch: the tile's first position its stream owns
    if (ns) {
        __syncthreads();
        const int si = s_str;
        first = si >= 0 && sbeg[si] > t0 ? sbeg[si] : t0;
        out = si >= 0 ? (uint8_t*)(sdst[si] - sbeg[si]) : nullptr;  // indexed by position: out + p
        if (si < 0) first = t0 + CRC_TILE;  // no stream: nothing written
    }
    for (uint32_t c = tid; c < CRC_TILE / 8; c += 256) {
        // positions 8c .. 8c+7 of the tile's frame; position k holds output byte t0 + k - pad
        uint32_t v[8];
        if (!pad) {
            const uint4 q = *(const uint4*)(wf + t0 + 8 * c);
            v[0] = q.x & 0xFFFF, v[1] = q.x >> 16, v[2] = q.y & 0xFFFF, v[3] = q.y >> 16;
            v[4] = q.z & 0xFFFF, v[5] = q.z >> 16, v[6] = q.w & 0xFFFF, v[7] = q.w >> 16;
        } else {
            for (int i = 0; i < 8; i++) {
                const uint32_t k = 8 * c + i;
                v[i] = k >= pad ? wf[t0 + k - pad] : 0u;
            }
        }
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t b = v[i] < 256 ? v[i] : ws[v[i] - 256];
            if (i < 4)
                lo |= b << (8 * i);
            else
                hi |= b << (8 * (i - 4));
        }
        if (ns) {
            const uint64_t p = t0 + 8 * c;
            if (p >= first) {
                *(U2*)(out + p) = make_uint2(lo, hi);
            } else if (p + 8 > first) {
                for (int i = 0; i < 8; i++)
                    if (p + i >= first) out[p + i] = (uint8_t)((i < 4 ? lo >> (8 * i) : hi >> (8 * (i - 4))) & 255);
            }
        } else if (!pad) {
            *(U2*)(out + t0 + 8 * c) = make_uint2(lo, hi);
        } else {
            for (int i = 0; i < 8; i++) {
                const uint32_t k = 8 * c + i;
                if (k >= pad) out[t0 + k - pad] = (uint8_t)((i < 4 ? lo >> (8 * i) : hi >> (8 * (i - 4))) & 255);
            }
        }
        const uint32_t w = 2 * c;  // word index in the tile frame
        rows[(w / (TD_PER / 4)) * TD_ROW + w % (TD_PER / 4)] = lo;
        rows[((w + 1) / (TD_PER / 4)) * TD_ROW + (w + 1) % (TD_PER / 4)] = hi;
    }
    __syncthreads();
    uint32_t r = 0;
    const uint32_t* row = rows + tid * TD_ROW;
#pragma unroll 4
    for (uint32_t i = 0; i < TD_PER / 4; i++) {
        const uint32_t x = r ^ row[i];
        r = tab[3][x & 255] ^ tab[2][(x >> 8) & 255] ^ tab[1][(x >> 16) & 255] ^ tab[0][x >> 24];
    }
    // the tile's register: every thread's part shifted past the bytes after it (shifts[t] =
    // x^(8 TD_PER (255 - t)) mod P), all XORed together (the register is linear in the data)
    r = gf2_mul(r, shifts[tid]);
#pragma unroll
    for (int d = 32; d; d >>= 1) r ^= __shfl_xor(r, d, 64);
    if ((tid & 63) == 0) part[tid >> 6] = r;
    __syncthreads();
    if (tid == 0) tile_crc[blockIdx.x] = part[0] ^ part[1] ^ part[2] ^ part[3];
}

// ------------------------------------------------------------------ the sink's thread
// Output reaches the caller's sink from one library thread, in stream order, one call at a time
// (the interface's rule). The pinned output buffers are a ring of OUT_BUFS slots with a length
// each (0: empty): the decoding side fills the slot at fill_ once the sink thread has emptied
// it, and the sink thread empties the slot at drain_ once it holds bytes. After a sink returns
// nonzero the remaining slots are emptied without being shown to it.
class SinkPump {
  public:
    SinkPump(gi_sink sink, void* user, uint8_t* const* slots)
        : sink_(sink), user_(user), slots_(slots), th_([this] { drain_loop(); }) {}
    SinkPump(const SinkPump&) = delete;
    SinkPump& operator=(const SinkPump&) = delete;
    ~SinkPump() { close(); }

    // the slot to fill next, as soon as it is empty
    int acquire() {
        std::unique_lock<std::mutex> lk(mu_);
        moved_.wait(lk, [&] { return len_[fill_] == 0; });
        return fill_;
    }
    // the acquired slot now holds n > 0 bytes
    void publish(size_t n) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            len_[fill_] = n;
            fill_ = (fill_ + 1) % OUT_BUFS;
        }
        moved_.notify_all();
    }
    bool stopped() const { return stopped_.load(std::memory_order_acquire); }
    // waits until every published slot is emptied; true when the sink stopped the delivery
    bool close() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            closing_ = true;
        }
        moved_.notify_all();
        if (th_.joinable()) th_.join();
        return stopped();
    }

  private:
    void drain_loop() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            moved_.wait(lk, [&] { return closing_ || len_[drain_] != 0; });
            const size_t n = len_[drain_];
            if (n == 0) return;  // closing with nothing left to show
            lk.unlock();
            if (!stopped() && sink_ && sink_(user_, slots_[drain_], n) != 0) stopped_.store(true, std::memory_order_release);
            lk.lock();
            len_[drain_] = 0;
            drain_ = (drain_ + 1) % OUT_BUFS;
            moved_.notify_all();
        }
    }

    const gi_sink sink_;
    void* const user_;
    uint8_t* const* const slots_;
    std::mutex mu_;
    std::condition_variable moved_;  // a slot was filled or emptied, or closing began
    size_t len_[OUT_BUFS] = {};
    int fill_ = 0, drain_ = 0;
    bool closing_ = false;
    std::atomic<bool> stopped_{false};
    std::thread th_;  // declared last: it starts once the members above exist
};

// ------------------------------------------------------------------ per-kernel time
// HIP events around each launch on the context's stream, from a fixed pool; flushed (and the
// times added up per kernel) whenever the host waits for the stream anyway.
enum KernelId { KID_FIND, KID_VERIFY, KID_TABLES, KID_PROBE, KID_DECODE, KID_GROUPS, KID_STORED, KID_PIECES,
                KID_TILE_A, KID_WIN_INIT, KID_WIN_JUMP, KID_WIN_BYTES, KID_TILE_D, KID_FIND_GRAPH, KID_RESOLVE_GRAPH,
                KID_COUNT };
// find_graph: the finder's kernels (k_find, k_verify, k_tables) as one graph launch;
// resolve_graph: the resolve's (k_stored, k_pieces, k_tile_a, k_win_*, k_tile_d)
const char* const KERNEL_NAMES[KID_COUNT] = {"k_find",     "k_verify",    "k_tables", "k_probe",
                                             "k_decode",   "k_groups",      "k_stored",    "k_pieces", "k_tile_a",
                                             "k_win_init", "k_win_jump",    "k_win_bytes", "k_tile_d", "find_graph",
                                             "resolve_graph"};

struct KernelClock {
    static constexpr int POOL = 512;
    hipEvent_t ev[2 * POOL] = {};
    int id[POOL] = {};
    int used = 0;
    uint64_t launches[KID_COUNT] = {};
    double seconds[KID_COUNT] = {};
    hipStream_t st = nullptr;

    void create(hipStream_t s) {
        st = s;
        for (auto& e : ev) HIP_OK(hipEventCreate(&e));
    }
    void destroy() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
    }
    void reset() {
        used = 0;
        for (int k = 0; k < KID_COUNT; k++) launches[k] = 0, seconds[k] = 0;
    }
    void begin(int k) {
        if (used == POOL) flush();
        id[used] = k;
        HIP_OK(hipEventRecord(ev[2 * used], st));
    }
    void end() {
        HIP_OK(hipEventRecord(ev[2 * used + 1], st));
        used++;
    }
    void flush() {
        if (!used) return;
        HIP_OK(hipEventSynchronize(ev[2 * used - 1]));
        for (int i = 0; i < used; i++) {
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
            launches[id[i]]++;
            seconds[id[i]] += ms / 1e3;
        }
        used = 0;
    }
};

}  // namespace

// ------------------------------------------------------------------ context

struct gi_ctx {
    int device = 0;
    std::string name, err;
    bool debug = false;
    // a batch worker's work since its batch's last reset (gi_batch_kernel_times): the kernel
    // clock's counts of each call added up, batch segments and streams decoded alone
    std::mutex acc_mu;
    uint64_t acc_launches[KID_COUNT] = {};
    double acc_seconds[KID_COUNT] = {};
    uint64_t acc_segments = 0, acc_alone = 0;
    void accumulate(bool segment) {
        std::lock_guard<std::mutex> lk(acc_mu);
        for (int k = 0; k < KID_COUNT; k++) {
            acc_launches[k] += kc.launches[k] + kcf.launches[k];
            acc_seconds[k] += kc.seconds[k] + kcf.seconds[k];
        }
        (segment ? acc_segments : acc_alone)++;
    }
    // GI_FAULT_AFTER_SEGMENTS=N when the context was made: each of its calls fails with
    // GI_ERR_GPU as it would begin segment N + 1 (-1: never), so that a caller's fallback to the
    // CPU can be tested on demand
    int fault_after = -1;
    Params P;
    uint64_t in_cap = 0;
    unsigned maxcand = 0, maxhits = 0, maxstored = 0, maxtables = 0, half_tables = 0;
    uint64_t max_super = 0;
    uint64_t ncu = 256;  // compute units
    uint64_t probe_lead = PROBE_LEAD;
    hipStream_t st = nullptr;
    hipEvent_t ev[12] = {};  // stage marks 1-5 of a segment, two sets (segments alternate)
    hipEvent_t up_ev[2] = {};  // the two halves of the input staging buffer: their last copy
    // the input written on st (an upload, a batch's packing): the finder's stream waits for it
    // before its next search, which reads the input from another stream
    hipEvent_t in_ev = nullptr;
    bool in_wait = false;
    void input_written() {
        if (sf == st) return;
        HIP_OK(hipEventRecord(in_ev, st));
        in_wait = true;
    }
    uint32_t last_crc = 0;
    KernelClock kc;

    uint8_t* d_in = nullptr;
    uint32_t* d_ests = nullptr;  // the verified headers' ratio estimates (k_tables)
    uint64_t *d_cand = nullptr, *d_hits = nullptr, *d_keys = nullptr, *d_newkeys = nullptr, *d_off = nullptr;
    uint64_t *d_poff = nullptr, *d_pdst = nullptr;
    unsigned *d_ncand = nullptr, *d_nhits = nullptr, *d_nstored = nullptr, *d_nscand = nullptr;
    unsigned* d_gqueue = nullptr;  // k_groups' work queue head
    unsigned groups_grid = 0;      // k_groups' persistent launch: its workgroups (0: one per queue item)
    uint64_t* d_scand = nullptr;  // stored-block candidates
    unsigned maxscand = 0;
    uint32_t *d_slots = nullptr, *d_newslots = nullptr, *d_cap = nullptr, *d_plen = nullptr;
    uint32_t *d_crc_tab = nullptr, *d_shifts = nullptr, *d_tile_crc = nullptr;
    uint32_t* d_wref = nullptr;  // pass B's window entries
    ConstMul shift_tile;         // multiplies a CRC register by x^(8 CRC_TILE)
    uint16_t* d_wf = nullptr;    // window form of the segment's output
    uint8_t* d_wins = nullptr;   // window bytes before every super tile
    uint8_t* d_wtmp = nullptr;
    int *d_tfirst = nullptr, *d_order = nullptr;
    Table* d_tabs = nullptr;
    Probe* d_probes = nullptr;
    Start *d_probed = nullptr, *d_jobs = nullptr, *d_targets = nullptr;
    Result* d_res = nullptr;
    GroupDesc* d_groups = nullptr;
    GroupResult* d_gres = nullptr;
    GroupHead* d_ghead = nullptr;  // the groups' results as the host reads them
    // pinned, allocated with the context (never while decoding: pinned allocation and release
    // synchronise the device, which can wait on other threads' pageable copies): a decode round's
    // results on their way to the host, for rounds of up to dl_cap jobs (larger ones copy to
    // pageable memory)
    Result* h_res_dl = nullptr;
    GroupHead* h_ghead = nullptr;
    size_t dl_cap = 0;
    double sync_wait = 0;  // seconds the host spent waiting for the streams (debug report)
    void sync() {
        Stopwatch t;
        HIP_OK(hipStreamSynchronize(st));
        sync_wait += t.seconds();
        stage_off = 0;  // every upload from the staging ring is done
    }
    // pinned staging for host-to-device uploads on st: copies of host data that the host may
    // change before the stream reaches them
    uint8_t* h_stage = nullptr;
    uint64_t stage_cap = 0, stage_off = 0;
    template <class T>
    const T* stage(const T* src, size_t n) {
        const uint64_t bytes = (n * sizeof(T) + 255) / 256 * 256;
        if (bytes > stage_cap) return src;  // too large: copied from where it is (a pageable copy, staged by the runtime before it returns)
        if (stage_off + bytes > stage_cap) sync();
        uint8_t* at = h_stage + stage_off;
        std::memcpy(at, src, n * sizeof(T));
        stage_off += bytes;
        return (const T*)at;
    }
    uint32_t* h_tile_crc[2] = {};  // pinned: the tiles' CRC registers of the last two segments
    void sync_find() { HIP_OK(hipStreamSynchronize(sf)); }
    hipStream_t sf = nullptr;  // the block finder's stream
    bool sf_shared = false;
    bool find_graphs = true;             // the finder's launches as one graph a segment
    hipGraphExec_t find_exec = nullptr;  // that graph, updated in place from segment to segment
    hipGraphExec_t resolve_exec = nullptr;  // the resolve's kernels as one graph, the same way
    uint64_t max_pieces = 0;
    ChainEntry* d_ents = nullptr;  // the chain entries of a segment (device, pinned host)
    ChainEntry* h_ents = nullptr;
    KernelClock kcf;           // its kernels' times
    unsigned* h_cnt = nullptr; // pinned: the finder's counts
    bool trace_groups = false;
    uint64_t* d_trace = nullptr;
    hipEvent_t tev[2] = {};
    uint16_t* d_out16 = nullptr;
    StoredCopy* d_stored = nullptr;
    // batches: the jobs' input ends, the packing chunks, the streams' output map (device, pinned)
    uint64_t* d_lim = nullptr;
    PackChunk *d_pack = nullptr, *h_pack = nullptr;
    uint64_t max_pack = 0;
    uint64_t *d_sbeg = nullptr, *d_sdst = nullptr, *h_sbeg = nullptr, *h_sdst = nullptr;
    uint8_t *d_final = nullptr, *d_window = nullptr;
    uint8_t* h_out[OUT_BUFS] = {};
    uint8_t* h_in = nullptr;
    size_t max_tiles = 0;

    std::vector<void*> dev_allocs, host_allocs;

    template <class T>
    void dalloc(T*& p, uint64_t count) {
        void* v = nullptr;
        HIP_OK(hipMalloc(&v, std::max<uint64_t>(count, 1) * sizeof(T)));
        dev_allocs.push_back(v);
        p = (T*)v;
    }
    template <class T>
    void halloc(T*& p, uint64_t count) {
        void* v = nullptr;
        HIP_OK(hipHostMalloc(&v, std::max<uint64_t>(count, 1) * sizeof(T), hipHostMallocDefault));
        host_allocs.push_back(v);
        p = (T*)v;
    }
    void release() {
        for (void* p : dev_allocs) (void)hipFree(p);
        for (void* p : host_allocs) (void)hipHostFree(p);
        dev_allocs.clear();
        host_allocs.clear();
        if (d_in) (void)hipFree(d_in);
        d_in = nullptr;
        in_cap = 0;
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        for (auto& e : tev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        for (auto& e : up_ev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        if (in_ev) (void)hipEventDestroy(in_ev), in_ev = nullptr;
        kc.destroy();
        kcf.destroy();
        if (find_exec) (void)hipGraphExecDestroy(find_exec), find_exec = nullptr;
        if (resolve_exec) (void)hipGraphExecDestroy(resolve_exec), resolve_exec = nullptr;
        if (st) (void)hipStreamDestroy(st), st = nullptr;
        if (sf && !sf_shared) (void)hipStreamDestroy(sf);
        sf = nullptr;
    }
    // The input buffer is made with the context, sized for a segment's search (so a batch
    // segment, which packs at most that much, never reallocates it on a worker thread beside
    // other contexts' work). Only a single-stream call on the caller's thread (gi_inflate,
    // gi_rescue_headers) with a longer stream replaces it, while nothing of this context runs.
    static uint64_t input_bytes(uint64_t n) { return (n + INPUT_PAD + 4095) / 4096 * 4096; }
    void alloc_input(uint64_t n) {
        void* v = nullptr;
        HIP_OK(hipMalloc(&v, input_bytes(n)));
        d_in = (uint8_t*)v;
        in_cap = input_bytes(n);
    }
    void ensure_input(uint64_t n) {
        if (input_bytes(n) <= in_cap) return;
        sync_find();
        sync();
        HIP_OK(hipFree(d_in));
        d_in = nullptr;
        in_cap = 0;
        alloc_input(n);
    }
};

namespace {

// The schedule's fields of a gi_config_tuned, checked against their ranges (0: the default)
struct Tuning {
    uint32_t chunk_kib = 0, probe_spacing_kbit = 0, probe_lead_bits = 0, seg_in_max_mb = 0, lanes_per_group = 0,
             persistent_per_cu = 0;
};

// segment_in_kb of gi_config: 0 (segments sized from the stream's ratio) or a forced size from
// 16 KiB (the smallest the reference's own tests force, gpuinflate/README.md:55-56) up to the
// context's search buffers (1 GiB of input a segment)
const char* segment_in_error(uint32_t kb) {
    if (kb && (kb < 16 || kb > (1u << 20))) return "segment_in_kb outside 16..1048576";
    return nullptr;
}

const char* tuning_error(const Tuning& t) {
    auto bad = [](uint32_t v, uint32_t lo, uint32_t hi) { return v && (v < lo || v > hi); };
    // chunks from the reference's 4 KiB stride up to 64 KiB; probes no further apart than the
    // 128 KiB at which a group's lanes still split a stretch; leads from 1 Kibit (short leads
    // synchronise less often, never wrongly: a false start is never landed on); input segments
    // from 64 MiB up to the 1 GiB the context's search buffers hold
    if (bad(t.chunk_kib, 4, 64)) return "chunk_kib outside 4..64";
    if (bad(t.probe_spacing_kbit, 64, 1024)) return "probe_spacing_kbit outside 64..1024";
    if (bad(t.probe_lead_bits, 1024, 8192)) return "probe_lead_bits outside 1024..8192";
    if (bad(t.seg_in_max_mb, 64, 1024)) return "segment_in_max_mb outside 64..1024";
    if (t.lanes_per_group && !group_lanes_ok((int)t.lanes_per_group)) return "lanes_per_group not 16, 32 or 64";
    if (bad(t.persistent_per_cu, 1, 16)) return "persistent_per_cu outside 1..16";
    return nullptr;
}

void setup(gi_ctx& g, const gi_config& cfg, const Tuning& tune) {
    g.device = cfg.device;
    g.debug = getenv("GI_DEBUG") != nullptr;
    if (const char* f = getenv("GI_FAULT_AFTER_SEGMENTS")) g.fault_after = std::max(0, atoi(f));
    Stopwatch t0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw GpuFault("no HIP device");
    if (cfg.device < 0 || cfg.device >= ndev) throw GpuFault("no such device");
    HIP_OK(hipSetDevice(g.device));
    HIP_OK(hipFree(nullptr));
    double t_ctx = t0.seconds();
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, g.device));
    g.name = prop.name[0] ? prop.name : prop.gcnArchName;
    if (g.name.empty()) g.name = "AMD GPU";
    HIP_OK(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking));
    // GZ_ONE_STREAM: the finder shares the decoders' stream (kernel times without overlap)
    if (getenv("GZ_ONE_STREAM"))
        g.sf = g.st, g.sf_shared = true;
    else
        HIP_OK(hipStreamCreateWithFlags(&g.sf, hipStreamNonBlocking));
    for (auto& e : g.ev) HIP_OK(hipEventCreate(&e));
    for (auto& e : g.up_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIP_OK(hipEventCreateWithFlags(&g.in_ev, hipEventDisableTiming));
    g.kc.create(g.st);
    g.kcf.create(g.sf);
    // GZ_NO_GRAPH: the finder's kernels launched one by one, each timed by the kernel clock
    g.find_graphs = getenv("GZ_NO_GRAPH") == nullptr;

    uint64_t seg_in_max = (uint64_t)(tune.seg_in_max_mb ? tune.seg_in_max_mb : 1024) << 20;
    if (cfg.segment_out_mb)  // a context for smaller streams: its buffers sized to match
        seg_in_max = std::min<uint64_t>(seg_in_max, std::max<uint64_t>(64ull << 20, 2ull * cfg.segment_out_mb << 20));
    g.P.seg_in_max = seg_in_max;
    const uint64_t nchunks_max = seg_in_max * 8 / CHUNK_BITS + 2;
    g.P.max_jobs = (size_t)(4 * nchunks_max + 8192);
    g.P.max_probes = g.P.max_jobs / 2;
    g.maxcand = (unsigned)std::min<uint64_t>(seg_in_max * 8 / 256, 1u << 26);
    // verified headers: about one per 20 KiB of level-6 data; room for one per 4 KiB
    g.maxhits = (unsigned)std::clamp<uint64_t>(seg_in_max / 4096, 16384, 1u << 18);
    g.half_tables = g.maxhits + std::max(4096u, g.maxhits / 4);  // a segment's tables: its headers and a few more
    g.maxtables = 1 + 2 * g.half_tables;  // slot 0: the fixed code; then two halves, one a segment
    g.maxstored = (unsigned)(seg_in_max / 1024 + 65536);
    g.probe_lead = tune.probe_lead_bits ? tune.probe_lead_bits : PROBE_LEAD;
    g.P.probe_spacing =
        tune.probe_spacing_kbit ? (uint64_t)tune.probe_spacing_kbit << 10 : PROBE_SPACING;
    // a coarser stride than CHUNK_BITS only lowers the job count the buffers are sized for
    g.P.chunk_bits = tune.chunk_kib ? (uint64_t)tune.chunk_kib * 8192 : CHUNK_BITS;
    g.P.group_lanes = tune.lanes_per_group ? (int)tune.lanes_per_group : GROUP_LANES_DEFAULT;
    uint64_t fixed = (uint64_t)g.maxcand * 8 + (uint64_t)g.maxhits * 8 + (uint64_t)g.maxtables * (sizeof(Table) + 24) +
                     (uint64_t)g.maxstored * sizeof(StoredCopy) +
                     (uint64_t)g.P.max_jobs * (3 * sizeof(Start) + sizeof(Result) + sizeof(Probe) + 64) + (256ull << 20);
    if (cfg.segment_out_mb) {
        g.P.seg_out_max = std::min<uint64_t>((uint64_t)cfg.segment_out_mb << 20, SEG_OUT_CAP);
    } else {
        size_t free_b = 0, total_b = 0;
        HIP_OK(hipMemGetInfo(&free_b, &total_b));
        double usable = 0.5 * (double)free_b - (double)fixed;
        uint64_t s = usable > 0 ? (uint64_t)(usable / 9.0) >> 20 << 20 : 0;
        g.P.seg_out_max = std::min<uint64_t>(s, 4096ull << 20);
        g.P.seg_out_max = std::min<uint64_t>(g.P.seg_out_max, SEG_OUT_CAP);
    }
    if (g.P.seg_out_max < SEG_OUT_MIN) throw GpuFault("not enough free GPU memory");
    g.P.forced_seg_in = (uint64_t)cfg.segment_in_kb << 10;
    // the decoders' output: 2.5 symbols per output byte of a segment (speculative overlap and the
    // room each decoder keeps), and at least room for a few thousand groups at their smallest
    g.P.pool = std::max<uint64_t>(g.P.seg_out_max / 2 * 5, 4096ull * g.P.group_lanes * (JOB_CAP_MIN / 4));
    g.P.debug = g.debug;
    g.ncu = (uint64_t)std::max(1, prop.multiProcessorCount);
    {
        int per_cu = 0;
        const void* k = g.P.group_lanes == 16   ? (const void*)k_groups<16, true>
                        : g.P.group_lanes == 64 ? (const void*)k_groups<64, true>
                                                : (const void*)k_groups<32, true>;
        HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0));
        const int want = (int)tune.persistent_per_cu;
        g.groups_grid = want ? (unsigned)(std::max(1, std::min(want, per_cu)) * g.ncu) : 0;
    }
    g.max_super = std::max<uint64_t>(std::max<uint64_t>(SUPER_SPLIT, g.ncu), (g.P.seg_out_max + SUPER_MAX - 1) / SUPER_MAX) + 2;
    g.max_tiles = (size_t)((g.P.seg_out_max + CRC_TILE - 1) / CRC_TILE + 1);

    Stopwatch t1;
    g.dalloc(g.d_cand, g.maxcand);
    g.dalloc(g.d_hits, g.maxhits);
    g.dalloc(g.d_ests, g.maxhits);
    g.dalloc(g.d_ncand, 1);
    g.maxscand = (unsigned)std::min<uint64_t>(seg_in_max / 2048 + 65536, 1u << 26);
    g.dalloc(g.d_scand, g.maxscand);
    g.dalloc(g.d_nscand, 1);
    g.dalloc(g.d_nhits, 1);
    g.dalloc(g.d_tabs, g.maxtables);
    g.dalloc(g.d_keys, g.maxtables);
    g.dalloc(g.d_slots, g.maxtables);
    g.dalloc(g.d_newkeys, g.maxtables);
    g.dalloc(g.d_newslots, g.maxtables);
    g.dalloc(g.d_probes, g.P.max_probes);
    g.dalloc(g.d_probed, g.P.max_probes);
    g.dalloc(g.d_jobs, g.P.max_jobs);
    g.dalloc(g.d_targets, g.P.max_jobs);
    g.dalloc(g.d_tfirst, g.P.max_jobs);
    g.dalloc(g.d_order, g.P.max_jobs);
    g.dalloc(g.d_off, g.P.max_jobs);
    g.dalloc(g.d_cap, g.P.max_jobs);
    g.dalloc(g.d_res, g.P.max_jobs);
    g.dalloc(g.d_groups, g.P.max_jobs);
    g.dalloc(g.d_gres, g.P.max_jobs);
    g.dalloc(g.d_ghead, g.P.max_jobs);
    g.trace_groups = getenv("GZ_TRACE_GROUPS") != nullptr;
    if (g.trace_groups) {
        g.dalloc(g.d_trace, (4 + 3 * (uint64_t)g.P.group_lanes) * g.P.max_jobs);
        for (auto& e : g.tev) HIP_OK(hipEventCreate(&e));
    }
    g.max_pieces = (uint64_t)g.P.max_jobs * g.P.group_lanes;
    g.dalloc(g.d_poff, g.max_pieces);
    g.dalloc(g.d_pdst, g.max_pieces);
    g.dalloc(g.d_plen, g.max_pieces);
    g.dalloc(g.d_ents, g.P.max_jobs + BATCH_MAX_STREAMS);
    g.dalloc(g.d_lim, g.P.max_jobs);
    g.max_pack = seg_in_max / PACK_CHUNK + BATCH_MAX_STREAMS + 16;
    g.dalloc(g.d_pack, g.max_pack);
    g.dalloc(g.d_sbeg, BATCH_MAX_STREAMS);
    g.dalloc(g.d_sdst, BATCH_MAX_STREAMS);
    g.dalloc(g.d_stored, g.maxstored);
    g.dalloc(g.d_nstored, 1);
    g.dalloc(g.d_gqueue, 1);
    g.dalloc(g.d_out16, g.P.pool);
    g.dalloc(g.d_wf, g.P.seg_out_max);
    g.dalloc(g.d_wins, g.max_super * WINDOW);
    g.dalloc(g.d_wtmp, WINDOW);
    g.dalloc(g.d_final, g.P.seg_out_max);
    g.dalloc(g.d_window, WINDOW);
    g.dalloc(g.d_crc_tab, 1024);
    g.dalloc(g.d_wref, g.max_super * WINDOW);
    g.dalloc(g.d_shifts, 256);
    g.dalloc(g.d_tile_crc, g.max_tiles);
    g.alloc_input(seg_in_max);
    double t_dev = t1.seconds();
    Stopwatch t2;
    for (auto& b : g.h_out) g.halloc(b, OUT_BUF);
    g.halloc(g.h_cnt, 4);
    g.halloc(g.h_ents, g.P.max_jobs + BATCH_MAX_STREAMS);
    g.dl_cap = std::min<size_t>(g.P.max_jobs, 131072);
    g.halloc(g.h_res_dl, g.dl_cap);
    g.halloc(g.h_ghead, g.dl_cap);
    g.stage_cap = 64ull << 20;
    g.halloc(g.h_stage, g.stage_cap);
    for (auto& t : g.h_tile_crc) g.halloc(t, g.max_tiles);
    g.halloc(g.h_pack, g.max_pack);
    g.halloc(g.h_sbeg, BATCH_MAX_STREAMS);
    g.halloc(g.h_sdst, BATCH_MAX_STREAMS);
    g.halloc(g.h_in, IN_BUF);
    // constants: the CRC table, the tree's shifts, the fixed code in slot 0
    // slice-by-4 tables: tab[k][n] is the CRC register after byte n and k zero bytes
    uint32_t tab[4][256];
    for (uint32_t n = 0; n < 256; n++) {
        uint32_t c = n;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        tab[0][n] = c;
    }
    for (int k = 1; k < 4; k++)
        for (uint32_t n = 0; n < 256; n++) tab[k][n] = (tab[k - 1][n] >> 8) ^ tab[0][tab[k - 1][n] & 255];
    g.shift_tile.set(x8n(CRC_TILE));
    uint32_t shifts[256];  // pass D: a thread's 64-byte part shifted past the parts after it
    for (int t = 0; t < 256; t++) shifts[t] = x8n((uint64_t)TD_PER * (255 - t));
    HIP_OK(hipMemcpy(g.d_crc_tab, tab, sizeof tab, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(g.d_shifts, shifts, sizeof shifts, hipMemcpyHostToDevice));
    uint64_t fk = UINT64_MAX;
    uint32_t fs = 0;
    HIP_OK(hipMemcpy(g.d_newkeys, &fk, 8, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(g.d_newslots, &fs, 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tables, dim3(1), dim3(64), 0, g.st, (const uint32_t*)nullptr, (uint64_t)0,
                       (const uint64_t*)g.d_newkeys, (const uint32_t*)g.d_newslots, 1, (const unsigned*)nullptr, 0u, 0u,
                       g.d_tabs, (uint32_t*)nullptr);
    HIP_OK(hipGetLastError());
    g.sync();
    if (g.debug)
        fprintf(stderr, "  setup: context %.0f ms, GPU memory %.0f ms, pinned memory %.0f ms (segments up to %llu MiB out)\n",
                t_ctx * 1e3, t_dev * 1e3, t2.seconds() * 1e3, (unsigned long long)(g.P.seg_out_max >> 20));
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, a, b));
    return ms / 1e3f;
}

// The HIP side of plan.h's inflate_stream
struct HipBackend {
    // find_wait may run on another thread beside the decode (it touches only the finder's stream
    // and buffers)
    static constexpr bool async_find = true;
    gi_ctx& g;
    const uint8_t* d_in;   // the stream in device memory
    uint64_t size;
    uint8_t* d_user_out;   // device-resident output
    uint64_t out_cap;
    bool device_mode;      // output stays in d_user_out; else it goes to the sink
    SinkPump* pump;
    uint64_t out_before = 0;
    uint32_t crc_raw = 0;  // raw CRC register of the output so far
    uint64_t crc_len = 0;
    bool stored_reset = false;
    double up_s = 0;

    const uint32_t* words() const { return (const uint32_t*)d_in; }

    bool aborted() { return pump && pump->stopped(); }
    bool finish_delivery() { return pump ? pump->close() : false; }

    // Stage marks: segments alternate between two sets of events, so that the host never waits
    // for a segment's resolve to read its times; they are read a segment later (the next decode
    // has waited for the stream by then), the last ones by settle()
    int par = 0;
    bool times_pending[2] = {false, false};
    hipEvent_t ev(int i) const { return g.ev[par * 6 + i]; }
    int segments_begun = 0;
    void begin_segment() {
        // the fault hook: the context's calls fail here as a device fault would
        if (g.fault_after >= 0 && segments_begun++ >= g.fault_after)
            throw GpuFault("injected fault after " + std::to_string(g.fault_after) + " segments (GI_FAULT_AFTER_SEGMENTS)");
        par ^= 1;
        HIP_OK(hipEventRecord(ev(1), g.st));
        stored_reset = false;
    }
    void mark(int i) { HIP_OK(hipEventRecord(ev(i), g.st)); }
    void read_times(int p, Stats& s) {
        hipEvent_t* e = g.ev + p * 6;
        HIP_OK(hipEventSynchronize(e[5]));  // done already, unless that segment decoded nothing
        s.find_s += elapsed(e[1], e[2]);
        s.decode_s += elapsed(e[2], e[3]);
        s.resolve_s += elapsed(e[3], e[4]);
        s.download_s += elapsed(e[4], e[5]);
        times_pending[p] = false;
    }
    void add_stage_times(Stats& s) {
        if (times_pending[par ^ 1]) read_times(par ^ 1, s);
        times_pending[par] = true;
        if (g.debug) fprintf(stderr, "    host waited for the GPU %.1f ms in the segment\n", g.sync_wait * 1e3);
        g.sync_wait = 0;
    }
    // the end of a call: the stream drained, every deferred CRC and time read
    void settle(Stats& s) {
        g.sync();
        fold_crc();
        for (int p = 0; p < 2; p++)
            if (times_pending[p]) read_times(p, s);
    }

    // the block finder over bits (lo, hi), on the finder's own stream: it runs beside the work
    // of the segment before (the two share only the input)
    void find_launch(uint64_t lo, uint64_t hi) {
        find_lo = lo;
        find_hi = hi;
        find_empty = hi <= lo + 1;
        find_half ^= 1;  // the half of the table cache the segment before is not using
        if (find_empty) return;
        hipStream_t sf = g.sf;
        // the input came on the decoders' stream: the search waits for it (outside the graph)
        if (g.in_wait) {
            HIP_OK(hipStreamWaitEvent(sf, g.in_ev, 0));
            g.in_wait = false;
        }
        if (!g.find_graphs) {
            find_enqueue(sf, true);
            return;
        }
        // the same sequence as one graph: captured, then the context's executable graph updated
        // in place (or instantiated again when the sequence's shape changed) and launched
        g.kcf.begin(KID_FIND_GRAPH);
        HIP_OK(hipStreamBeginCapture(sf, hipStreamCaptureModeThreadLocal));
        hipGraph_t graph = nullptr;
        try {
            find_enqueue(sf, false);
        } catch (...) {
            (void)hipStreamEndCapture(sf, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIP_OK(hipStreamEndCapture(sf, &graph));
        if (g.find_exec) {
            hipGraphNode_t bad_node = nullptr;
            hipGraphExecUpdateResult r;
            if (hipGraphExecUpdate(g.find_exec, graph, &bad_node, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(g.find_exec);
                g.find_exec = nullptr;
            }
        }
        hipError_t e = g.find_exec ? hipSuccess : hipGraphInstantiate(&g.find_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIP_OK(e);
        HIP_OK(hipGraphLaunch(g.find_exec, sf));
        g.kcf.end();
    }
    // the finder's sequence on sf: counters cleared by a kernel (no memset nodes in a graph),
    // the bit and stored-block searches, the header verifier, the headers' tables, and the two
    // counts copied to pinned memory. `timed`: each kernel between events of the kernel clock
    void find_enqueue(hipStream_t sf, bool timed) {
        const uint64_t lo = find_lo, hi = find_hi;
        auto t0 = [&](int k) { if (timed) g.kcf.begin(k); };
        auto t1 = [&] { if (timed) g.kcf.end(); };
        hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, sf, g.d_ncand, g.d_nhits, g.d_nscand);
        uint64_t word_lo = lo / 32, word_hi = (hi + 31) / 32;
        uint64_t nwords = word_hi - word_lo;
        // stored blocks: LEN at bytes from the segment start to its end (and the bytes a block
        // starting just before `hi` needs), in the same pass over the words
        const uint64_t byte_lo = lo / 8, byte_hi = std::min<uint64_t>(size, hi / 8 + 8);
        const uint64_t nwords_s = byte_hi > 4 * word_lo ? (byte_hi - 4 * word_lo + 3) / 4 : 0;
        t0(KID_FIND);
        const unsigned find_blocks =
            (unsigned)std::min<uint64_t>((std::max(nwords, nwords_s) + FIND_WORDS - 1) / FIND_WORDS, 4096);
        hipLaunchKernelGGL(k_find, dim3(find_blocks), dim3(256), 0, sf, words(), size, word_lo, nwords, lo, hi,
                           g.d_cand, g.d_ncand, g.maxcand, byte_lo, byte_hi, g.d_scand, g.d_nscand, g.maxscand);
        t1();
        HIP_OK(hipGetLastError());
        t0(KID_VERIFY);
        // grids sized to what the range can hold: candidates (under a quarter of its bits pass the
        // prefilter) and verified headers (one per 2^14 bits is many)
        const uint64_t span = hi - lo;
        const unsigned vblocks = (unsigned)std::min<uint64_t>(blocks_for(std::min<uint64_t>(g.maxcand, span / 4 + 256), 256), 4 * g.ncu);
        const unsigned tblocks = (unsigned)std::min<uint64_t>(span / 16384 + 64, 32768);
        hipLaunchKernelGGL(k_verify, dim3(vblocks), dim3(256), 0, sf, words(), size, (const uint64_t*)g.d_cand,
                           (const unsigned*)g.d_ncand, g.maxcand, g.d_hits, g.d_nhits, g.maxhits);
        t1();
        HIP_OK(hipGetLastError());
        // every verified header's table
        t0(KID_TABLES);
        hipLaunchKernelGGL(k_tables, dim3(tblocks), dim3(64), 0, sf, words(), size, (const uint64_t*)g.d_hits,
                           (const uint32_t*)nullptr, 0, (const unsigned*)g.d_nhits, g.maxhits, half_base(find_half),
                           g.d_tabs, g.d_ests);
        t1();
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(&g.h_cnt[0], g.d_nhits, sizeof(unsigned), hipMemcpyDeviceToHost, sf));
        HIP_OK(hipMemcpyAsync(&g.h_cnt[1], g.d_nscand, sizeof(unsigned), hipMemcpyDeviceToHost, sf));
    }
    uint32_t half_base(int h) const { return 1 + (uint32_t)h * g.half_tables; }
    int find_half = 1;
    Found find_wait() {
        Found out;
        out.slot_base = half_base(find_half);
        out.prebuilt = true;
        if (find_empty) return out;
        g.sync_find();
        const unsigned nh = std::min(g.h_cnt[0], g.maxhits), ns = std::min(g.h_cnt[1], g.maxscand);
        out.dyn.resize(nh);
        out.est.resize(nh);
        out.stored.resize(ns);
        if (nh) HIP_OK(hipMemcpyAsync(out.dyn.data(), g.d_hits, nh * 8ull, hipMemcpyDeviceToHost, g.sf));
        if (nh) HIP_OK(hipMemcpyAsync(out.est.data(), g.d_ests, nh * 4ull, hipMemcpyDeviceToHost, g.sf));
        if (ns) HIP_OK(hipMemcpyAsync(out.stored.data(), g.d_scand, ns * 8ull, hipMemcpyDeviceToHost, g.sf));
        g.sync_find();
        g.kcf.flush();
        return out;
    }
    uint64_t find_lo = 0, find_hi = 0;
    bool find_empty = true;

    void tables(const std::vector<std::pair<uint64_t, uint32_t>>& fresh, const TableKeys& all) {
        // this segment's slots run from its half's base to next_slot - 1
        if (all.next_slot > all.base + g.half_tables) throw GpuFault("table cache full");
        if (!fresh.empty()) {
            std::vector<uint64_t> k(fresh.size());
            std::vector<uint32_t> s(fresh.size());
            for (size_t i = 0; i < fresh.size(); i++) {
                k[i] = fresh[i].first;
                s[i] = fresh[i].second;
            }
            HIP_OK(hipMemcpyAsync(g.d_newkeys, g.stage(k.data(), k.size()), k.size() * 8, hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_newslots, g.stage(s.data(), s.size()), s.size() * 4, hipMemcpyHostToDevice, g.st));
            g.kc.begin(KID_TABLES);
            hipLaunchKernelGGL(k_tables, dim3((unsigned)std::min<size_t>(k.size(), 32768)), dim3(64), 0, g.st, words(), size,
                               (const uint64_t*)g.d_newkeys, (const uint32_t*)g.d_newslots, (int)k.size(),
                               (const unsigned*)nullptr, 0u, 0u, g.d_tabs, (uint32_t*)nullptr);
            g.kc.end();
            HIP_OK(hipGetLastError());
        }
        // through the pinned staging ring: no wait for the stream (a resolve may still run on it)
        if (!all.keys.empty()) {
            HIP_OK(hipMemcpyAsync(g.d_keys, g.stage(all.keys.data(), all.keys.size()), all.keys.size() * 8,
                              hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_slots, g.stage(all.slots.data(), all.slots.size()), all.slots.size() * 4,
                              hipMemcpyHostToDevice, g.st));
        }
        nkeys = (int)all.keys.size();
    }
    int nkeys = 0;

    std::vector<Start> probe(const std::vector<Probe>& pts, const TableKeys&) {
        std::vector<Start> out(pts.size());
        if (pts.empty()) return out;
        HIP_OK(hipMemcpyAsync(g.d_probes, pts.data(), pts.size() * sizeof(Probe), hipMemcpyHostToDevice, g.st));
        { g.kc.begin(KID_PROBE); hipLaunchKernelGGL(k_probe, dim3(blocks_for(pts.size(), 256)), dim3(256), 0, g.st, words(), size,
                           (const Table*)g.d_tabs, (const uint64_t*)g.d_keys, (const uint32_t*)g.d_slots, nkeys,
                           (const Probe*)g.d_probes, (int)pts.size(), g.probe_lead, g.d_probed); g.kc.end(); }
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out.data(), g.d_probed, pts.size() * sizeof(Start), hipMemcpyDeviceToHost, g.st));
        g.sync();
        return out;
    }

    std::vector<int> lanes_buf;
    std::vector<GroupDesc> gds_buf;
    void decode(Jobs& j, size_t a, size_t b, const TableKeys&, bool stream_start, uint64_t stop, bool first_round) {
        size_t n = b - a;
        if (!n) return;
        if (j.start.size() > g.P.max_jobs) throw GpuFault("too many jobs");
        // every job's region lies inside the pool (make_jobs gives each job a floor of room, which
        // a segment of tiny blocks could push past it)
        if (j.used > g.P.pool) throw GpuFault("the decoders' output needs more room than the pool holds");
        if (first_round) {
            HIP_OK(hipMemcpyAsync(g.d_targets, g.stage(j.targets.data(), j.targets.size()), j.targets.size() * sizeof(Start),
                              hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemsetAsync(g.d_nstored, 0, sizeof(unsigned), g.st));
        }
        std::vector<int>& lanes = lanes_buf;  // members: their memory is reused round to round
        std::vector<GroupDesc>& gds = gds_buf;
        lanes.clear();
        gds.clear();
        for (size_t k = a; k < b; k++) {
            if (j.kind[k] == JOB_GROUP) {
                GroupDesc d{};
                d.s = j.start[k];
                d.end = j.end[k];
                d.pool = j.off[k];
                d.lane_cap = j.cap[k];
                d.split_end = j.split_end[k];
                d.hard = j.hard[k];
                d.last_extra = j.extra[k];
                d.tfirst = j.tfirst[k];
                d.job = (int)k;
                d.lim = j.lim[k];
                gds.push_back(d);
            } else {
                lanes.push_back((int)k);
            }
        }
        // A later round's few lane jobs ride in the group launch, each a group of lane 0 alone
        // (lane_cap 0: no split, its whole room, every target after it), instead of a k_decode
        // launch of their own before it: the two would run one after the other, each mostly
        // waiting for its longest lane
        if (!first_round && !lanes.empty() && !gds.empty() && lanes.size() <= FOLD_LANES) {
            for (int k : lanes) {
                GroupDesc d{};
                d.s = j.start[k];
                d.end = j.start[k].bit;
                d.pool = j.off[k];
                d.lane_cap = 0;
                d.last_extra = j.cap[k];
                d.tfirst = j.tfirst[k];
                d.job = k;
                d.lim = j.lim[k];
                gds.push_back(d);
            }
            lanes.clear();
        }
        StoredList sl{g.d_stored, g.d_nstored, g.maxstored};
        if (!lanes.empty()) {
            HIP_OK(hipMemcpyAsync(g.d_order, lanes.data(), lanes.size() * 4, hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_jobs + a, j.start.data() + a, n * sizeof(Start), hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_tfirst + a, j.tfirst.data() + a, n * 4, hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_off + a, j.off.data() + a, n * 8, hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_cap + a, j.cap.data() + a, n * 4, hipMemcpyHostToDevice, g.st));
            const bool lims = !j.streams.empty();
            if (lims) HIP_OK(hipMemcpyAsync(g.d_lim + a, j.lim.data() + a, n * 8, hipMemcpyHostToDevice, g.st));
            if (g.trace_groups) HIP_OK(hipEventRecord(g.tev[0], g.st));
            g.kc.begin(KID_DECODE);
            hipLaunchKernelGGL(k_decode, dim3(blocks_for(lanes.size(), 256)), dim3(256), 0, g.st, words(), size,
                               (const Table*)g.d_tabs, (const uint64_t*)g.d_keys, (const uint32_t*)g.d_slots, nkeys,
                               (const Start*)g.d_jobs, (const int*)g.d_tfirst, (const uint64_t*)g.d_off,
                               (const uint32_t*)g.d_cap, (const Start*)g.d_targets, (int)j.targets.size(),
                               (const int*)g.d_order, 0, (int)lanes.size(), g.d_out16, g.d_res, stream_start ? 1 : 0,
                               stop, sl, lims ? (const uint64_t*)g.d_lim : (const uint64_t*)nullptr);
            g.kc.end();
            HIP_OK(hipGetLastError());
            if (g.trace_groups) {
                HIP_OK(hipEventRecord(g.tev[1], g.st));
                HIP_OK(hipEventSynchronize(g.tev[1]));
                std::vector<Result> rr(n);
                HIP_OK(hipMemcpy(rr.data(), g.d_res + a, n * sizeof(Result), hipMemcpyDeviceToHost));
                uint32_t mx = 0;
                for (int k : lanes) mx = std::max(mx, rr[k - a].out_len);
                fprintf(stderr, "    k_decode: %zu lanes, %.0f us, longest %u symbols\n", lanes.size(),
                        elapsed(g.tev[0], g.tev[1]) * 1e6, mx);
            }
        }
        // the groups' results go by job index: the jobs [g0, g1) hold them all
        size_t g0 = SIZE_MAX, g1 = 0;
        for (const GroupDesc& d : gds) g0 = std::min(g0, (size_t)d.job), g1 = std::max(g1, (size_t)d.job + 1);
        if (!gds.empty()) {
            // a persistent launch's queue, longest interval first (a long group taken last would
            // run alone); a workgroup per item keeps the planner's order (sorting measured no
            // faster there, and it costs the host a sort a round)
            if (g.groups_grid)
                std::stable_sort(gds.begin(), gds.end(), [](const GroupDesc& x, const GroupDesc& y) {
                    return x.end - x.s.bit > y.end - y.s.bit;
                });
            HIP_OK(hipMemcpyAsync(g.d_groups, g.stage(gds.data(), gds.size()), gds.size() * sizeof(GroupDesc),
                              hipMemcpyHostToDevice, g.st));
            if (g.groups_grid) HIP_OK(hipMemsetAsync(g.d_gqueue, 0, sizeof(unsigned), g.st));
            const int L = g.P.group_lanes;
            const bool loop = g.groups_grid != 0;
            const unsigned grid = (unsigned)std::min<uint64_t>(blocks_for(gds.size(), 64 / L), loop ? g.groups_grid : UINT32_MAX);
            auto launch = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(grid), dim3(64), 0, g.st, words(), size, (const Table*)g.d_tabs,
                                   (const uint64_t*)g.d_keys, (const uint32_t*)g.d_slots, nkeys,
                                   (const GroupDesc*)g.d_groups, (int)gds.size(), g.d_gqueue, (const Start*)g.d_targets,
                                   (int)j.targets.size(), g.d_out16, g.d_gres, g.d_ghead, stream_start ? 1 : 0, stop,
                                   g.probe_lead, sl, g.trace_groups ? g.d_trace : nullptr);
            };
            g.kc.begin(KID_GROUPS);
            if (L == 16)
                loop ? launch(k_groups<16, true>) : launch(k_groups<16, false>);
            else if (L == 64)
                loop ? launch(k_groups<64, true>) : launch(k_groups<64, false>);
            else
                loop ? launch(k_groups<32, true>) : launch(k_groups<32, false>);
            g.kc.end();
            HIP_OK(hipGetLastError());
        }
        // the results: lane jobs' Results and groups' heads (by job index, first group to last),
        // through the pinned buffers when the round fits them
        const bool pinned = n <= g.dl_cap;
        if (!lanes.empty())
            HIP_OK(hipMemcpyAsync(pinned ? g.h_res_dl : j.res.data() + a, g.d_res + a, n * sizeof(Result),
                              hipMemcpyDeviceToHost, g.st));
        if (!gds.empty()) {
            if (pinned)
                HIP_OK(hipMemcpyAsync(g.h_ghead, g.d_ghead + g0, (g1 - g0) * sizeof(GroupHead), hipMemcpyDeviceToHost, g.st));
            else
                HIP_OK(hipMemcpyAsync(j.ghead.data() + g0, g.d_ghead + g0, (g1 - g0) * sizeof(GroupHead),
                                  hipMemcpyDeviceToHost, g.st));
        }
        g.sync();
        fold_crc();  // the segment before: its resolve is done
        if (pinned)
            for (int k : lanes) j.res[k] = g.h_res_dl[k - a];
        for (size_t i = 0; i < gds.size(); i++) {
            const size_t k = (size_t)gds[i].job;
            if (pinned) j.ghead[k] = g.h_ghead[k - g0];
            j.res[k] = j.ghead[k].r;
        }
        if (g.trace_groups && !gds.empty()) print_group_trace(gds);
    }

    // GZ_TRACE_GROUPS: the spread of group times in the last k_groups launch and its slowest groups
    void print_group_trace(const std::vector<GroupDesc>& gds) {
        const int L = g.P.group_lanes, GPW = 64 / L;
        std::vector<uint64_t> t(gds.size() * 4 + gds.size() * L * 3);
        HIP_OK(hipMemcpy(t.data(), g.d_trace, t.size() * 8, hipMemcpyDeviceToHost));
        uint64_t t0 = UINT64_MAX, t2 = 0;
        std::vector<std::pair<uint64_t, size_t>> d(gds.size());
        for (size_t i = 0; i < gds.size(); i++) {
            t0 = std::min(t0, t[i * 4]);
            t2 = std::max(t2, t[i * 4 + 2]);
            d[i] = {t[i * 4 + 2] - t[i * 4], i};
        }
        std::sort(d.begin(), d.end());
        auto us = [](uint64_t ticks) { return ticks / 100.0; };
        fprintf(stderr, "    groups %zu: span %.0f us; group time p50 %.0f p90 %.0f p99 %.0f max %.0f us\n", gds.size(),
                us(t2 - t0), us(d[d.size() / 2].first), us(d[d.size() * 9 / 10].first), us(d[d.size() * 99 / 100].first),
                us(d.back().first));
        {
            // balance: per wavefront (GPW groups), its lanes' symbols against 64 x its longest lane's;
            // probe and decode time shares
            double sum_out = 0, sum_max = 0, tp = 0, td = 0;
            uint64_t lanes_run = 0, lanes_over = 0;
            for (size_t i0 = 0; i0 < gds.size(); i0 += GPW) {
                uint32_t mx = 0;
                double su = 0;
                for (size_t i = i0; i < gds.size() && i < i0 + GPW; i++) {
                    tp += (double)(t[i * 4 + 1] - t[i * 4]);
                    td += (double)(t[i * 4 + 2] - t[i * 4 + 1]);
                    const uint64_t* lt = t.data() + 4 * gds.size() + i * L * 3;
                    for (int l = 0; l < L; l++) {
                        const uint64_t z = lt[l * 3 + 2];
                        if (z == ~0ull) continue;
                        const uint32_t o = (uint32_t)(z & 0xFFFFFFFF);
                        su += o;
                        mx = std::max(mx, o);
                        lanes_run++;
                        lanes_over += (int)(int8_t)(z >> 32 & 0xFF) == ERR_OVERFLOW;
                    }
                }
                sum_out += su;
                sum_max += 64.0 * mx;
            }
            fprintf(stderr, "    balance: lanes' symbols %.0f, 64 x longest lane per wavefront %.0f (%.2f); probe %.0f%% of "
                            "group time; %llu lanes ran, %llu hit their cap\n",
                    sum_out, sum_max, sum_out / std::max(1.0, sum_max), 100.0 * tp / std::max(1.0, tp + td),
                    (unsigned long long)lanes_run, (unsigned long long)lanes_over);
        }
        for (size_t q = d.size() > 12 ? d.size() - 12 : 0; q < d.size(); q++) {
            size_t i = d[q].second;
            const GroupDesc& gd = gds[i];
            uint64_t x = t[i * 4 + 3];
            fprintf(stderr, "      job %d mode %d bits %llu split %llu cap %u extra %u: start %.0f probe %.0f decode %.0f us; "
                            "longest lane %u out %u status %d; %u lanes started\n",
                    gd.job, gd.s.mode, (unsigned long long)(gd.end - gd.s.bit), (unsigned long long)(gd.split_end - gd.s.bit),
                    gd.lane_cap, gd.last_extra, us(t[i * 4] - t0), us(t[i * 4 + 1] - t[i * 4]), us(t[i * 4 + 2] - t[i * 4 + 1]),
                    (unsigned)(x >> 32 & 0xFF), (unsigned)(x & 0xFFFFFFFF), (int)(int8_t)(x >> 40 & 0xFF), (unsigned)(x >> 48));
        }
        if (getenv("GZ_TRACE_LANES"))
            for (size_t i = d.back().second & ~(size_t)(GPW - 1); i < gds.size() && i < (d.back().second | (GPW - 1)) + 1; i++) {
            const uint64_t* lt = t.data() + 4 * gds.size() + i * L * 3;
            fprintf(stderr, "      slowest group, job %d from %llu to %llu (split %llu):\n", gds[i].job,
                    (unsigned long long)gds[i].s.bit, (unsigned long long)gds[i].end, (unsigned long long)gds[i].split_end);
            for (int l = 0; l < L; l++) {
                const uint64_t z = lt[l * 3 + 2];
                if (z == ~0ull) {
                    fprintf(stderr, "        lane %2d start %llu: not run\n", l, (unsigned long long)lt[l * 3]);
                    continue;
                }
                fprintf(stderr, "        lane %2d start %llu end %llu: status %d landed %d out %u\n", l,
                        (unsigned long long)lt[l * 3], (unsigned long long)lt[l * 3 + 1], (int)(int8_t)(z >> 32 & 0xFF),
                        (int)(int16_t)(z >> 40 & 0xFFFF), (unsigned)(z & 0xFFFFFFFF));
            }
        }
    }

    // where this segment's bytes go, and where the W bytes before them are
    uint8_t* seg_dst() const { return device_mode ? d_user_out + out_before : g.d_final; }
    const uint8_t* window_src(uint32_t W) const { return device_mode ? d_user_out + out_before - W : g.d_window; }

    // Pieces to bytes: stored copies, pass A, pass B (the W window bytes before the segment at
    // win), pass D into dst (or, ns > 0, the batch's streams mapped in d_sbeg / d_sdst); returns
    // every CRC_TILE's raw CRC register in tc
    // tc null: the CRC registers go to pinned memory and are combined into the stream's CRC later
    // (fold_crc), so that the host need not wait for the resolve
    void resolve_core(const Pieces& p, uint32_t W, const uint8_t* win, uint8_t* dst, int ns, std::vector<uint32_t>* tc) {
        const uint64_t total = p.total;
        if (tc) tc->clear();
        if (p.npieces > g.max_pieces || p.jobs.size() > g.P.max_jobs + BATCH_MAX_STREAMS)
            throw GpuFault("more pieces than the buffers hold (internal error)");
        const uint64_t super = total ? super_size(total, g.ncu) : 0;
        const unsigned nsuper = total ? (unsigned)((total + super - 1) / super) : 0;
        if (nsuper > g.max_super) throw GpuFault("segment larger than its buffers (internal error)");
        if (total) {
            // the entries go up through pinned memory (the copy does not wait for the stream)
            std::memcpy(g.h_ents, p.jobs.data(), p.jobs.size() * sizeof(ChainEntry));
            HIP_OK(hipMemcpyAsync(g.d_ents, g.h_ents, p.jobs.size() * sizeof(ChainEntry), hipMemcpyHostToDevice, g.st));
        }
        if (!g.find_graphs || g.trace_groups) {
            resolve_enqueue(p, W, win, dst, ns, super, nsuper, true);
        } else {
            // the same kernels as one graph: captured, the context's executable graph updated in
            // place (instantiated again when the sequence's shape changed: the number of doubling
            // rounds, or an empty segment) and launched
            g.kc.begin(KID_RESOLVE_GRAPH);
            HIP_OK(hipStreamBeginCapture(g.st, hipStreamCaptureModeThreadLocal));
            hipGraph_t graph = nullptr;
            try {
                resolve_enqueue(p, W, win, dst, ns, super, nsuper, false);
            } catch (...) {
                (void)hipStreamEndCapture(g.st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            HIP_OK(hipStreamEndCapture(g.st, &graph));
            if (g.resolve_exec) {
                hipGraphNode_t bad_node = nullptr;
                hipGraphExecUpdateResult r;
                if (hipGraphExecUpdate(g.resolve_exec, graph, &bad_node, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(g.resolve_exec);
                    g.resolve_exec = nullptr;
                }
            }
            hipError_t e = g.resolve_exec ? hipSuccess : hipGraphInstantiate(&g.resolve_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_OK(e);
            HIP_OK(hipGraphLaunch(g.resolve_exec, g.st));
            g.kc.end();
        }
        if (!total) return;
        const uint64_t tiles = (total + CRC_TILE - 1) / CRC_TILE;
        // the tiles' raw CRC registers
        if (tc) {
            tc->resize(tiles);
            HIP_OK(hipMemcpyAsync(tc->data(), g.d_tile_crc, tiles * 4, hipMemcpyDeviceToHost, g.st));
            g.sync();
        } else {
            if (crc_pend) {  // two segments' registers in flight: settle the older
                g.sync();
                fold_crc();
            }
            crc_buf ^= 1;
            HIP_OK(hipMemcpyAsync(g.h_tile_crc[crc_buf], g.d_tile_crc, tiles * 4, hipMemcpyDeviceToHost, g.st));
            crc_pend = true;
            crc_tiles = tiles;
            crc_total = total;
        }
    }

    // the resolve's kernels on the decode stream (no copies, no waits: a graph can hold them).
    // `timed`: each kernel between events of the kernel clock
    void resolve_enqueue(const Pieces& p, uint32_t W, const uint8_t* win, uint8_t* dst, int ns, uint64_t super,
                         unsigned nsuper, bool timed) {
        auto t0 = [&](int k) { if (timed) g.kc.begin(k); };
        auto t1 = [&] { if (timed) g.kc.end(); };
        const uint64_t total = p.total;
        t0(KID_STORED);
        hipLaunchKernelGGL(k_stored, dim3((unsigned)std::min<uint64_t>(size / 16384 + 16, 4096)), dim3(256), 0, g.st,
                           (const StoredCopy*)g.d_stored,
                           (const unsigned*)g.d_nstored, g.maxstored, d_in, g.d_out16);
        t1();
        HIP_OK(hipGetLastError());
        if (!total) return;
        const int np = (int)p.npieces;
        t0(KID_PIECES);
        hipLaunchKernelGGL(k_pieces, dim3(blocks_for(p.jobs.size(), 256)), dim3(256), 0, g.st,
                           (const ChainEntry*)g.d_ents, (int)p.jobs.size(), (const GroupResult*)g.d_gres, g.d_poff,
                           g.d_plen, g.d_pdst);
        t1();
        HIP_OK(hipGetLastError());
        t0(KID_TILE_A);
        hipLaunchKernelGGL(k_tile_a, dim3(nsuper), dim3(TA_THREADS), 0, g.st, (const uint16_t*)g.d_out16,
                           (const uint64_t*)g.d_poff, (const uint32_t*)g.d_plen, (const uint64_t*)g.d_pdst, np, total,
                           super, g.d_wf, g.trace_groups ? g.d_trace : nullptr);
        t1();
        HIP_OK(hipGetLastError());
        if (g.trace_groups) {
            std::vector<uint64_t> t(5ull * nsuper);
            g.sync();
            HIP_OK(hipMemcpy(t.data(), g.d_trace, t.size() * 8, hipMemcpyDeviceToHost));
            double a = 0, b = 0, c = 0, r = 0, f = 0;
            for (unsigned i = 0; i < nsuper; i++)
                a += t[5 * i], b += t[5 * i + 1], c += t[5 * i + 2], r += t[5 * i + 3], f += t[5 * i + 4];
            const double nt = (double)nsuper * (double)(super / TILE);  // tiles, 10 ns ticks
            fprintf(stderr, "    k_tile_a: %u super tiles of %llu KiB, per tile: pieces %.2f us, fill %.2f us, chains %.2f us (%.1f rounds), writes %.2f us\n",
                    nsuper, (unsigned long long)(super >> 10), a / nt / 100.0, f / nt / 100.0, b / nt / 100.0, r / nt, c / nt / 100.0);
        }
        // pass B: the window bytes before every super tile
        const unsigned wblocks = (unsigned)((uint64_t)nsuper * WINDOW / 256);
        t0(KID_WIN_INIT);
        hipLaunchKernelGGL(k_win_init, dim3(wblocks), dim3(256), 0, g.st, (const uint16_t*)g.d_wf, win, W,
                           (int)nsuper, super, g.d_wref);
        t1();
        int rounds = 1;
        while ((1u << (rounds - 1)) < nsuper) rounds++;
        for (int r = 0; r < rounds && nsuper > 2; r++) {
            t0(KID_WIN_JUMP);
            hipLaunchKernelGGL(k_win_jump, dim3(wblocks), dim3(256), 0, g.st, (int)nsuper, g.d_wref);
            t1();
        }
        t0(KID_WIN_BYTES);
        hipLaunchKernelGGL(k_win_bytes, dim3((wblocks + 3) / 4), dim3(256), 0, g.st, (int)nsuper,
                           (const uint32_t*)g.d_wref, g.d_wins);
        t1();
        HIP_OK(hipGetLastError());
        const uint64_t tiles = (total + CRC_TILE - 1) / CRC_TILE;
        t0(KID_TILE_D);
        hipLaunchKernelGGL(k_tile_d, dim3((unsigned)tiles), dim3(256), 0, g.st, (const uint16_t*)g.d_wf, total, super,
                           (const uint8_t*)g.d_wins, dst, (const uint32_t*)g.d_crc_tab, (const uint32_t*)g.d_shifts,
                           g.d_tile_crc, (const uint64_t*)g.d_sbeg, (const uint64_t*)g.d_sdst, ns);
        t1();
        HIP_OK(hipGetLastError());
    }

    int resolve(const Jobs&, const Pieces& p, uint32_t W, uint64_t ob, std::string& err) {
        out_before = ob;
        const uint64_t total = p.total;
        if (device_mode && out_before + total > out_cap) {
            err = "the output is larger than the buffer given";
            return RC_ARG;
        }
        uint8_t* dst = seg_dst();
        resolve_core(p, W, window_src(W), dst, 0, nullptr);
        // the window for the next segment (host delivery keeps it apart from d_final)
        if (!device_mode) {
            const uint64_t newW = std::min<uint64_t>(WINDOW, (uint64_t)W + total);
            if (total >= newW) {
                HIP_OK(hipMemcpyAsync(g.d_window, dst + total - newW, newW, hipMemcpyDeviceToDevice, g.st));
            } else {
                const uint64_t keep = newW - total;  // bytes of the old window still in reach
                HIP_OK(hipMemcpyAsync(g.d_wtmp, g.d_window + (W - keep), keep, hipMemcpyDeviceToDevice, g.st));
                if (total) HIP_OK(hipMemcpyAsync(g.d_wtmp + keep, dst, total, hipMemcpyDeviceToDevice, g.st));
                HIP_OK(hipMemcpyAsync(g.d_window, g.d_wtmp, newW, hipMemcpyDeviceToDevice, g.st));
            }
        }
        return RC_OK;
    }

    // a batch: every finished stream's bytes to its own buffer (out_ptr[i]) and its CRC-32
    std::vector<uint64_t> out_ptr;
    void resolve_batch(const Jobs&, const Pieces& p, std::vector<BatchStream>& bs) {
        if (p.total % CRC_TILE) throw GpuFault("a batch's output is not in whole tiles (internal error)");
        int ns = 0;
        for (size_t i = 0; i < bs.size(); i++) {
            if (bs[i].rc != RC_OK || !bs[i].out_size) continue;
            if (bs[i].out_size > bs[i].out_cap) throw GpuFault("batch output past its buffer (internal error)");
            g.h_sbeg[ns] = bs[i].out_off;
            g.h_sdst[ns] = out_ptr[i];
            ns++;
        }
        if (ns) {
            HIP_OK(hipMemcpyAsync(g.d_sbeg, g.h_sbeg, ns * 8ull, hipMemcpyHostToDevice, g.st));
            HIP_OK(hipMemcpyAsync(g.d_sdst, g.h_sdst, ns * 8ull, hipMemcpyHostToDevice, g.st));
        }
        std::vector<uint32_t> tc;
        resolve_core(p, 0, g.d_window, nullptr, ns, &tc);
        for (auto& b : bs) {
            if (b.rc != RC_OK) continue;
            uint32_t raw = 0;
            for (uint64_t t = b.out_off / CRC_TILE; t < (b.out_off + b.out_size) / CRC_TILE; t++)
                raw = g.shift_tile.mul(raw) ^ tc[t];
            b.crc = ~(raw ^ gf2_mul(x8n(b.out_size), 0xFFFFFFFFu));
        }
        HIP_OK(hipEventRecord(g.ev[5], g.st));
    }

    bool deliver(uint64_t total, uint32_t) {
        if (device_mode) {
            HIP_OK(hipEventRecord(g.ev[5], g.st));
            return true;
        }
        const uint8_t* src = g.d_final;
        for (uint64_t at = 0; at < total; at += OUT_BUF) {
            const uint64_t len = std::min<uint64_t>(OUT_BUF, total - at);
            const int slot = pump->acquire();
            HIP_OK(hipMemcpyAsync(g.h_out[slot], src + at, len, hipMemcpyDeviceToHost, g.st));
            g.sync();
            pump->publish((size_t)len);
            if (pump->stopped()) return false;
        }
        HIP_OK(hipEventRecord(g.ev[5], g.st));
        return true;
    }

    // the segment whose tile CRCs are in flight (a pinned buffer of two)
    bool crc_pend = false;
    int crc_buf = 0;
    uint64_t crc_tiles = 0, crc_total = 0;
    // call with the stream past that segment's copy
    void fold_crc() {
        if (!crc_pend) return;
        const uint32_t* tc = g.h_tile_crc[crc_buf];
        for (uint64_t t = 0; t < crc_tiles; t++) {
            const uint64_t len = std::min<uint64_t>(CRC_TILE, crc_total - t * CRC_TILE);
            crc_raw = (len == CRC_TILE ? g.shift_tile.mul(crc_raw) : gf2_mul(crc_raw, x8n(len))) ^ tc[t];
        }
        crc_len += crc_total;
        crc_pend = false;
    }
    uint32_t crc32() const { return ~(crc_raw ^ gf2_mul(x8n(crc_len), 0xFFFFFFFFu)); }
};

int run_inflate(gi_ctx& g, const uint8_t* d_in, uint64_t size, uint64_t hint, bool device_mode, uint8_t* d_out,
                uint64_t out_cap, gi_sink sink, void* user, Stats& stats) {
    Stopwatch t0;
    HIP_OK(hipSetDevice(g.device));
    g.kc.flush();
    g.kc.reset();
    g.kcf.flush();
    g.kcf.reset();
    HipBackend be{g, d_in, size, d_out, out_cap, device_mode, nullptr};
    std::unique_ptr<SinkPump> pump;
    if (!device_mode) {
        pump = std::make_unique<SinkPump>(sink, user, g.h_out);
        be.pump = pump.get();
    }
    int rc;
    try {
        rc = inflate_stream(be, g.P, hint, size, stats, g.err);
        be.settle(stats);
    } catch (...) {
        (void)hipStreamSynchronize(g.sf);  // a finder run ahead reads the caller's input
        throw;
    }
    g.sync_find();
    if (pump) {
        const bool ab = pump->close();
        if (rc == RC_OK && ab) {
            g.err = "stopped by the caller";
            rc = RC_ABORTED;
        }
    }
    if (rc == RC_OK) g.last_crc = be.crc32();
    g.kc.flush();
    g.kcf.flush();
    stats.seconds = t0.seconds();
    return rc;
}

void to_gi_stats(const Stats& s, gi_stats* o) {
    if (!o) return;
    *o = gi_stats{s.in_used,  s.out_size, s.segments,  s.rounds,    s.jobs,       s.seconds,
                  s.upload_s, s.find_s,   s.decode_s, s.resolve_s, s.download_s};
}

// Runs one entry point's work on context g and turns what it throws into the interface's codes:
// a HIP failure or a buffer limit is GI_ERR_GPU (the caller decodes on the CPU), host memory
// running out too. After a failure both of the context's streams are drained and the runtime's
// sticky error is read, so the context's next call starts clean.
template <class Body>
int on_context(gi_ctx& g, Body&& body) {
    const char* what = nullptr;
    std::string msg;
    try {
        return body();
    } catch (const GpuFault& e) {
        msg = e.what();
    } catch (const std::bad_alloc&) {
        what = "out of host memory";
    }
    g.err = what ? what : msg;
    if (g.sf && g.sf != g.st) (void)hipStreamSynchronize(g.sf);
    if (g.st) (void)hipStreamSynchronize(g.st);
    (void)hipGetLastError();
    return GI_ERR_GPU;
}

// The checks every decoding entry point makes on its input before touching the device: a
// pointer with a size is required, and an empty stream is a truncated one (the reference's
// codes and messages, gpuinflate.cu:915-920)
int reject_input(gi_ctx& g, const void* in, uint64_t size) {
    if (size == 0) {
        g.err = "empty input";
        return GI_ERR_TRUNCATED;
    }
    if (in == nullptr) {
        g.err = "no input";
        return GI_ERR_ARG;
    }
    return GI_OK;
}

// The host's input copied to d_in through the two halves of the pinned staging buffer, so the
// host fills one half while the other is on its way; then the zero padding the decoders read
// past the end (only the padding: bytes beyond it are never read)
void upload_input(gi_ctx& g, const uint8_t* in, uint64_t n) {
    constexpr uint64_t HALF = IN_BUF / 2;
    for (uint64_t at = 0, k = 0; at < n; at += HALF, k++) {
        const int h = (int)(k & 1);
        if (k >= 2) HIP_OK(hipEventSynchronize(g.up_ev[h]));  // that half's previous copy has left
        const uint64_t m = std::min(HALF, n - at);
        std::memcpy(g.h_in + h * HALF, in + at, (size_t)m);
        HIP_OK(hipMemcpyAsync(g.d_in + at, g.h_in + h * HALF, m, hipMemcpyHostToDevice, g.st));
        HIP_OK(hipEventRecord(g.up_ev[h], g.st));
    }
    HIP_OK(hipMemsetAsync(g.d_in + n, 0, std::min<uint64_t>(g.in_cap - n, INPUT_PAD), g.st));
    g.input_written();
}

// A caller's configuration read into the library's: the reference's 16 bytes (a NULL pointer
// keeps every default, a size field under 16 is an error, gpuinflate.cu:888-892) and, when the
// size field says so, the schedule that follows them. Returns null or what is wrong.
const char* read_config(const gi_config* cfg, gi_config& base, Tuning& tune) {
    base = gi_config{sizeof(gi_config), 0, 0, 0};
    tune = Tuning{};
    if (cfg == nullptr) return nullptr;
    if (cfg->size < sizeof(gi_config)) return "gi_config too small";
    std::memcpy(&base, cfg, sizeof base);
    if (const char* bad = segment_in_error(base.segment_in_kb)) return bad;
    // a gi_config_tuned of 32 bytes (without lanes_per_group) or of its whole size
    if (cfg->size >= offsetof(gi_config_tuned, lanes_per_group)) {
        gi_config_tuned t{};
        std::memcpy(&t, cfg, std::min<size_t>(cfg->size, sizeof t));
        tune = {t.chunk_kib,       t.probe_spacing_kbit, t.probe_lead_bits,
                t.segment_in_max_mb, t.lanes_per_group,  t.persistent_per_cu};
        return tuning_error(tune);
    }
    return nullptr;
}

void copy_message(char* dst, size_t cap, const std::string& m) {
    if (dst && cap) snprintf(dst, cap, "%s", m.c_str());
}

}  // namespace

static_assert(sizeof(gi_config) == 16, "gi_config is 16 bytes");
static_assert(sizeof(gi_config_tuned) == 40 && offsetof(gi_config_tuned, lanes_per_group) == 32,
              "gi_config_tuned is 40 bytes");
static_assert(sizeof(gi_stats) == 80, "gi_stats is 80 bytes");

extern "C" {

GI_API uint32_t gi_abi_version(void) { return GI_ABI_VERSION; }

GI_API int gi_create(const gi_config* cfg, gi_ctx** out, char* err, size_t err_len) {
    if (out == nullptr) {
        copy_message(err, err_len, "no context pointer");
        return GI_ERR_ARG;
    }
    *out = nullptr;
    gi_config base;
    Tuning tune;
    if (const char* bad = read_config(cfg, base, tune)) {
        copy_message(err, err_len, bad);
        return GI_ERR_ARG;
    }
    auto g = std::make_unique<gi_ctx>();
    std::string why;
    try {
        setup(*g, base, tune);
    } catch (const GpuFault& e) {
        why = e.what();
    } catch (const std::bad_alloc&) {
        why = "out of host memory";
    }
    if (!why.empty()) {
        g->release();
        copy_message(err, err_len, why);
        return GI_ERR_GPU;
    }
    *out = g.release();
    return GI_OK;
}

GI_API int gi_inflate(gi_ctx* g, const uint8_t* in, uint64_t in_size, uint64_t out_hint, gi_sink sink, void* user,
                      gi_stats* stats) {
    if (g == nullptr) return GI_ERR_ARG;
    g->err.clear();
    Stats s;
    int rc = reject_input(*g, in, in_size);
    if (rc == GI_OK)
        rc = on_context(*g, [&] {
            HIP_OK(hipSetDevice(g->device));
            const Stopwatch up;
            g->ensure_input(in_size);
            upload_input(*g, in, in_size);
            const double up_s = up.seconds();
            const int r = run_inflate(*g, g->d_in, in_size, out_hint, false, nullptr, 0, sink, user, s);
            s.upload_s += up_s;
            return r;
        });
    to_gi_stats(s, stats);
    return rc;
}

GI_API int gi_inflate_device(gi_ctx* g, const void* d_in, uint64_t in_size, void* d_out, uint64_t out_cap,
                             uint64_t out_hint, uint64_t* out_size, uint32_t* crc32, gi_stats* stats) {
    if (g == nullptr) return GI_ERR_ARG;
    g->err.clear();
    Stats s;
    int rc = reject_input(*g, d_in, in_size);
    if (rc == GI_OK && d_out == nullptr && out_cap) {
        g->err = "no output buffer";
        rc = GI_ERR_ARG;
    }
    if (rc == GI_OK)
        rc = on_context(*g, [&] {
            return run_inflate(*g, (const uint8_t*)d_in, in_size, out_hint, true, (uint8_t*)d_out, out_cap, nullptr,
                               nullptr, s);
        });
    const bool ok = rc == GI_OK;
    if (out_size) *out_size = ok ? s.out_size : 0;
    if (crc32) *crc32 = ok ? g->last_crc : 0;
    to_gi_stats(s, stats);
    return rc;
}

// The rescue's header search (rescue.cpp:597-626, Rescuer::find_headers) on the GPU: every
// confirmed dynamic-block header in [from, 8 * size - 128) of in[0..size), in bit order
GI_API int gi_rescue_headers(gi_ctx* g, const uint8_t* in, uint64_t size, uint64_t from, uint64_t* out, uint64_t cap,
                             uint64_t* count) {
    if (!g || (!in && size) || (!out && cap) || !count || from == 0) return GI_ERR_ARG;
    g->err.clear();
    *count = 0;
    if (size > g->P.seg_in_max) {
        g->err = "input larger than a segment's search holds";
        return GI_ERR_ARG;
    }
    const uint64_t nbits = size * 8;
    const uint64_t end = nbits > 128 ? nbits - 128 : 0;
    if (from >= end) return GI_OK;
    return on_context(*g, [&] {
        gi_ctx& c = *g;
        HIP_OK(hipSetDevice(c.device));
        c.kc.flush();
        c.kc.reset();
        c.kcf.flush();
        c.kcf.reset();
        c.ensure_input(size);
        upload_input(c, in, size);
        HipBackend be{c, c.d_in, size, nullptr, 0, true, nullptr};
        // every plausible header from `from` to the end (the decodes' tables), searched as one
        // segment; the candidates are those before `end`
        be.find_launch(from ? from - 1 : 0, nbits);
        Found f = be.find_wait();
        TableKeys keys;
        keys.clear(f.slot_base);
        keys.prebuilt(f.dyn, f.slot_base);
        be.tables({}, keys);
        std::vector<uint64_t> cand;
        for (uint64_t b : f.dyn)
            if (b >= from && b < end) cand.push_back(b);
        std::sort(cand.begin(), cand.end());
        if (cand.size() > c.maxtables) throw GpuFault("more candidates than the buffers hold (internal error)");
        std::vector<uint32_t> ok(cand.size());
        if (!cand.empty()) {
            HIP_OK(hipMemcpyAsync(c.d_newkeys, cand.data(), cand.size() * 8, hipMemcpyHostToDevice, c.st));
            c.kc.begin(KID_DECODE);
            hipLaunchKernelGGL(k_rescue_verify, dim3(blocks_for(cand.size(), 64)), dim3(64), 0, c.st, be.words(), size,
                               (const Table*)c.d_tabs, (const uint64_t*)c.d_keys, (const uint32_t*)c.d_slots, be.nkeys,
                               (const uint64_t*)c.d_newkeys, (int)cand.size(), c.d_newslots);
            c.kc.end();
            HIP_OK(hipGetLastError());
            HIP_OK(hipMemcpyAsync(ok.data(), c.d_newslots, cand.size() * 4, hipMemcpyDeviceToHost, c.st));
        }
        c.sync();
        c.kc.flush();
        c.kcf.flush();
        uint64_t k = 0;
        for (size_t i = 0; i < cand.size(); i++)
            if (ok[i]) {
                if (k < cap) out[k] = cand[i];
                k++;
            }
        *count = k;
        return GI_OK;
    });
}

GI_API uint32_t gi_last_crc32(const gi_ctx* g) { return g ? g->last_crc : 0; }

GI_API int gi_kernel_times(const gi_ctx* g, gi_kernel_time* out, int max) {
    if (!g || !out) return 0;
    int n = 0;
    for (int k = 0; k < KID_COUNT && n < max; k++) {
        if (!g->kc.launches[k] && !g->kcf.launches[k]) continue;
        out[n].name = KERNEL_NAMES[k];
        out[n].launches = g->kc.launches[k] + g->kcf.launches[k];
        out[n].seconds = g->kc.seconds[k] + g->kcf.seconds[k];
        n++;
    }
    return n;
}

GI_API const char* gi_last_error(const gi_ctx* g) {
    static const char* const NONE = "no context";
    return g == nullptr ? NONE : g->err.c_str();
}

GI_API const char* gi_device_name(const gi_ctx* g) { return g == nullptr ? "" : g->name.c_str(); }

GI_API void gi_destroy(gi_ctx* g) {
    if (g == nullptr) return;
    (void)hipSetDevice(g->device);
    // nothing of the context's may still be queued when its memory goes
    if (g->sf && g->sf != g->st) (void)hipStreamSynchronize(g->sf);
    if (g->st) (void)hipStreamSynchronize(g->st);
    g->release();
    delete g;
}

}  // extern "C"

// ------------------------------------------------------------------ header search without allocation
// gi_rescue_headers_noalloc: the rescue's header search with every device buffer carved from one
// file-scope array, for callers that cannot allocate. Its host code uses the array's address as
// hipGetSymbolAddress gives it (once per device, with that device current), never the symbol.
namespace noalloc {

constexpr uint64_t INPUT_MAX = 16ull << 20;                       // input bytes at most
constexpr unsigned MAXCAND = (unsigned)(INPUT_MAX * 8 / 256);     // the finder's candidates (as a context sizes them)
constexpr unsigned MAXHITS = 16384;                               // plausible headers
constexpr unsigned NTABLES = 1 + MAXHITS;                         // slot 0: the fixed code
constexpr uint64_t al(uint64_t x) { return (x + 255) / 256 * 256; }
constexpr uint64_t OFF_IN = 0;
constexpr uint64_t OFF_CAND = OFF_IN + al(INPUT_MAX + INPUT_PAD);
constexpr uint64_t OFF_HITS = OFF_CAND + al(8ull * MAXCAND);
constexpr uint64_t OFF_ESTS = OFF_HITS + al(8ull * MAXHITS);
constexpr uint64_t OFF_TABS = OFF_ESTS + al(4ull * MAXHITS);
constexpr uint64_t OFF_KEYS = OFF_TABS + al((uint64_t)sizeof(Table) * NTABLES);
constexpr uint64_t OFF_SLOTS = OFF_KEYS + al(8ull * NTABLES);
constexpr uint64_t OFF_VCAND = OFF_SLOTS + al(4ull * NTABLES);
constexpr uint64_t OFF_OK = OFF_VCAND + al(8ull * MAXHITS);
constexpr uint64_t OFF_CNT = OFF_OK + al(4ull * MAXHITS);        // counters; then the fixed code's key and slot
constexpr uint64_t BYTES = OFF_CNT + 256;
static_assert(BYTES <= (96ull << 20), "the no-allocation scratch stays under 96 MiB");

// the host side of a call: fixed arrays, one call at a time in the process
std::mutex mu;
void* dev_ptr[256] = {};  // per device: the scratch array's device address
uint64_t h_hits[MAXHITS], h_keys[MAXHITS], h_vcand[MAXHITS];
uint32_t h_idx[MAXHITS], h_slots[MAXHITS], h_ok[MAXHITS];

}  // namespace noalloc

static __device__ __align__(256) uint8_t scratch[noalloc::BYTES];

namespace {

// the scratch array's address on device dev (current), taken once
uint8_t* noalloc_scratch(int dev) {
    if (!noalloc::dev_ptr[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        noalloc::dev_ptr[dev] = p;
    }
    return (uint8_t*)noalloc::dev_ptr[dev];
}

// The search itself, on the null stream, with the device current. Returns GI_OK or GI_ERR_ARG
// (with why) when the file holds more candidates or headers than the scratch does.
int noalloc_search(int dev, const void* d_in, uint64_t size, uint64_t from, uint64_t* out, uint64_t cap,
                   uint64_t* count, std::string& why) {
    using namespace noalloc;
    uint8_t* S = noalloc_scratch(dev);
    const hipStream_t st = nullptr;
    const uint32_t* words = (const uint32_t*)(S + OFF_IN);
    uint64_t* cand = (uint64_t*)(S + OFF_CAND);
    uint64_t* hits = (uint64_t*)(S + OFF_HITS);
    uint32_t* ests = (uint32_t*)(S + OFF_ESTS);
    Table* tabs = (Table*)(S + OFF_TABS);
    uint64_t* keys = (uint64_t*)(S + OFF_KEYS);
    uint32_t* slots = (uint32_t*)(S + OFF_SLOTS);
    uint64_t* vcand = (uint64_t*)(S + OFF_VCAND);
    uint32_t* ok = (uint32_t*)(S + OFF_OK);
    unsigned* cnt = (unsigned*)(S + OFF_CNT);             // [0] candidates, [1] headers, [2] unused
    uint64_t* fixed_key = (uint64_t*)(S + OFF_CNT + 64);  // UINT64_MAX: the fixed code, into slot 0
    uint32_t* fixed_slot = (uint32_t*)(S + OFF_CNT + 128);
    int ncu = 0;
    HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const uint64_t nbits = size * 8, end = nbits > 128 ? nbits - 128 : 0;
    // the input and the zero padding its readers look ahead into (bytes past it are never read)
    HIP_OK(hipMemcpyAsync(S + OFF_IN, d_in, size, hipMemcpyDeviceToDevice, st));
    HIP_OK(hipMemsetAsync(S + OFF_IN + size, 0, INPUT_PAD, st));
    HIP_OK(hipMemsetAsync(fixed_key, 0xFF, 8, st));
    HIP_OK(hipMemsetAsync(fixed_slot, 0, 4, st));
    hipLaunchKernelGGL(k_tables, dim3(1), dim3(64), 0, st, (const uint32_t*)nullptr, (uint64_t)0, (const uint64_t*)fixed_key,
                       (const uint32_t*)fixed_slot, 1, (const unsigned*)nullptr, 0u, 0u, tabs, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_zero3, dim3(1), dim3(64), 0, st, cnt, cnt + 1, cnt + 2);
    // the finder over bit positions [from, nbits), as gi_rescue_headers runs it
    const uint64_t lo = from - 1, hi = nbits, word_lo = lo / 32, nwords = (hi + 31) / 32 - word_lo, span = hi - lo;
    hipLaunchKernelGGL(k_find, dim3((unsigned)std::min<uint64_t>((nwords + FIND_WORDS - 1) / FIND_WORDS, 4096)), dim3(256), 0,
                       st, words, size, word_lo, nwords, lo, hi, cand, cnt, MAXCAND, (uint64_t)0, (uint64_t)0,
                       (uint64_t*)nullptr, (unsigned*)nullptr, 0u);
    const unsigned vblocks =
        (unsigned)std::min<uint64_t>(blocks_for(std::min<uint64_t>(MAXCAND, span / 4 + 256), 256), 4 * (uint64_t)ncu);
    hipLaunchKernelGGL(k_verify, dim3(vblocks), dim3(256), 0, st, words, size, (const uint64_t*)cand, (const unsigned*)cnt,
                       MAXCAND, hits, cnt + 1, MAXHITS);
    hipLaunchKernelGGL(k_tables, dim3((unsigned)std::min<uint64_t>(span / 16384 + 64, 32768)), dim3(64), 0, st, words, size,
                       (const uint64_t*)hits, (const uint32_t*)nullptr, 0, (const unsigned*)(cnt + 1), MAXHITS, 1u, tabs,
                       ests);
    HIP_OK(hipGetLastError());
    unsigned counts[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(counts, cnt, sizeof counts, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (counts[0] > MAXCAND || counts[1] > MAXHITS) {
        why = "the input holds more header candidates than the no-allocation scratch does (" + std::to_string(MAXCAND) +
              " candidates, " + std::to_string(MAXHITS) + " headers at most)";
        return GI_ERR_ARG;
    }
    const unsigned nh = counts[1];
    // the headers' tables are in slots 1 + i (report order): keys sorted, with their slots
    if (nh) HIP_OK(hipMemcpy(h_hits, hits, nh * 8ull, hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < nh; i++) h_idx[i] = i;
    std::sort(h_idx, h_idx + nh, [](uint32_t a, uint32_t b) { return h_hits[a] < h_hits[b]; });
    unsigned nv = 0;
    for (unsigned i = 0; i < nh; i++) {
        h_keys[i] = h_hits[h_idx[i]];
        h_slots[i] = 1 + h_idx[i];
        if (h_keys[i] >= from && h_keys[i] < end) h_vcand[nv++] = h_keys[i];
    }
    uint64_t k = 0;
    if (nv) {
        HIP_OK(hipMemcpyAsync(keys, h_keys, nh * 8ull, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(slots, h_slots, nh * 4ull, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(vcand, h_vcand, nv * 8ull, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_rescue_verify, dim3(blocks_for(nv, 64)), dim3(64), 0, st, words, size, (const Table*)tabs,
                           (const uint64_t*)keys, (const uint32_t*)slots, (int)nh, (const uint64_t*)vcand, (int)nv, ok);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(h_ok, ok, nv * 4ull, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (unsigned i = 0; i < nv; i++)
            if (h_ok[i]) {
                if (k < cap) out[k] = h_vcand[i];
                k++;
            }
    }
    *count = k;
    return GI_OK;
}

}  // namespace

extern "C" {

GI_API int gi_rescue_headers_noalloc(int device, const void* d_in, uint64_t size, uint64_t from, uint64_t* out,
                                     uint64_t cap, uint64_t* count, char* err, size_t err_len) {
    auto refuse = [&](int code, const std::string& m) {
        copy_message(err, err_len, m);
        return code;
    };
    if (count == nullptr || (d_in == nullptr && size) || (out == nullptr && cap) || from == 0)
        return refuse(GI_ERR_ARG, "bad arguments");
    *count = 0;
    if (size > noalloc::INPUT_MAX)
        return refuse(GI_ERR_ARG, "input larger than the no-allocation scratch holds (" +
                                      std::to_string(noalloc::INPUT_MAX) + " bytes at most)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return refuse(GI_ERR_GPU, "no HIP device");
    if (device < 0 || device >= ndev || device >= 256) return refuse(GI_ERR_ARG, "no such device");
    copy_message(err, err_len, "");
    if (size * 8 <= 128 || from >= size * 8 - 128) return GI_OK;
    std::lock_guard<std::mutex> lk(noalloc::mu);
    int prev = -1;
    (void)hipGetDevice(&prev);
    std::string why;
    int rc;
    try {
        HIP_OK(hipSetDevice(device));
        rc = noalloc_search(device, d_in, size, from, out, cap, count, why);
    } catch (const GpuFault& e) {
        why = e.what();
        rc = GI_ERR_GPU;
        (void)hipStreamSynchronize(nullptr);
        (void)hipGetLastError();
    }
    if (prev >= 0) (void)hipSetDevice(prev);
    if (rc != GI_OK) {
        *count = 0;
        return refuse(rc, why);
    }
    return GI_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
// Worker threads, each with a decoder context of its own (its own HIP streams and buffers),
// taking streams from one queue: each takes everything queued that fits one batch segment and
// decodes it with one set of launches (a stream too large to share a segment goes alone).
// gi_batch_run queues a whole batch largest first; gi_batch_submit queues one stream in arrival
// order and gi_batch_wait hands finished ones back (a serving loop: what arrives while the
// workers are busy forms their next batch).
namespace {

// Decodes jobs[0..n) as one batch segment (plan.h inflate_batch): their inputs packed into the
// context's input buffer, every stage one set of launches for all of them. Jobs the batch does
// not finish (rounds or room ran out), and every job after a GPU error in the batch, are then
// decoded alone. Sets every job's results.
void decode_alone(gi_ctx* g, gi_job* j);

void run_batch(gi_ctx* g, gi_job* const* jobs, size_t n) {
    std::vector<BatchStream> bs(n);
    bool alone = false;
    int rc = on_context(*g, [&] {
        gi_ctx& c = *g;
        HIP_OK(hipSetDevice(c.device));
        c.kc.flush();
        c.kc.reset();
        c.kcf.flush();
        c.kcf.reset();
        uint64_t at = 0;
        for (size_t i = 0; i < n; i++) {
            bs[i].base = at;
            bs[i].size = jobs[i]->in_size;
            bs[i].out_cap = jobs[i]->out_cap;
            at = (at + bs[i].size + 15) / 16 * 16 + 16;  // 16 zero bytes at least between streams
        }
        const uint64_t total = at;
        // a worker takes at most a segment's input (gi_batch::work), which the context was made with
        if (gi_ctx::input_bytes(total) > c.in_cap) throw GpuFault("a batch segment's input exceeds its buffer (internal error)");
        uint64_t nch = 0;
        for (size_t i = 0; i < n; i++) {
            const uint64_t end = bs[i].base + bs[i].size;
            const uint64_t next = i + 1 < n ? bs[i + 1].base : total;
            uint64_t off = 0;
            do {
                if (nch >= c.max_pack) throw GpuFault("too many packing chunks (internal error)");
                const uint64_t len = std::min<uint64_t>(PACK_CHUNK, bs[i].size - off);
                const bool last = off + len >= bs[i].size;
                c.h_pack[nch++] = {(const uint8_t*)jobs[i]->d_in + off, bs[i].base + off, (uint32_t)len,
                                   (uint32_t)(last ? len + (next - end) : len)};
                off += len;
            } while (off < bs[i].size);
        }
        HIP_OK(hipMemcpyAsync(c.d_pack, c.h_pack, nch * sizeof(PackChunk), hipMemcpyHostToDevice, c.st));
        hipLaunchKernelGGL(k_pack, dim3((unsigned)nch), dim3(256), 0, c.st, (const PackChunk*)c.d_pack, c.d_in);
        c.input_written();
        HIP_OK(hipGetLastError());
        // the gaps' source: the pool's first BATCH_ZERO symbols (a lone decode may have used them)
        HIP_OK(hipMemsetAsync(c.d_out16, 0, BATCH_ZERO * sizeof(uint16_t), c.st));
        HipBackend be{c, c.d_in, total, nullptr, 0, true, nullptr};
        be.out_ptr.resize(n);
        for (size_t i = 0; i < n; i++) be.out_ptr[i] = (uint64_t)jobs[i]->d_out;
        Stats st;
        try {
            inflate_batch(be, c.P, bs, total, st);
            be.settle(st);
        } catch (...) {
            (void)hipStreamSynchronize(c.sf);
            throw;
        }
        c.sync_find();
        c.sync();
        c.kc.flush();
        c.kcf.flush();
        c.accumulate(true);
        return GI_OK;
    });
    if (rc != GI_OK) alone = true;  // a GPU error in the batch: each job on its own
    for (size_t i = 0; i < n; i++) {
        gi_job* j = jobs[i];
        if (alone || bs[i].rc < 0) {
            decode_alone(g, j);
            continue;
        }
        j->status = bs[i].rc;
        j->out_size = bs[i].rc == RC_OK ? bs[i].out_size : 0;
        j->in_used = bs[i].rc == RC_OK ? bs[i].in_used : 0;
        j->crc32 = bs[i].rc == RC_OK ? bs[i].crc : 0;
    }
}

}  // namespace

namespace {

// How gi_batch sizes jobs for a worker's batch segment: its input packed 16-byte aligned with 16
// zero bytes after it (run_batch), its output room in whole BATCH_ALIGN units; a job over a
// quarter of a segment's input or output goes alone, with a whole segment's parallelism
struct JobSizes {
    static uint64_t in_bytes(const gi_job* j) { return (j->in_size + 15) / 16 * 16 + 16; }
    static uint64_t out_bytes(const gi_job* j) { return (j->out_cap + BATCH_ALIGN - 1) / BATCH_ALIGN * BATCH_ALIGN; }
    static bool alone(const gi_job* j, const SegmentRoom& r) {
        return j->in_size > r.in_bytes / 4 || j->out_cap > r.out_bytes / 4;
    }
};
using JobQueue = BatchQueue<gi_job, JobSizes>;

// One job decoded by itself on context g, its results set
void decode_alone(gi_ctx* g, gi_job* j) {
    gi_stats st;
    uint64_t osz = 0;
    uint32_t crc = 0;
    j->status = gi_inflate_device(g, j->d_in, j->in_size, j->d_out, j->out_cap, j->out_hint, &osz, &crc, &st);
    g->accumulate(false);
    j->out_size = osz;
    j->crc32 = crc;
    j->in_used = j->status == GI_OK ? st.in_used : 0;
}

}  // namespace

struct gi_batch {
    std::vector<gi_ctx*> ctxs;
    std::vector<std::thread> threads;
    JobQueue queue;

    // a worker: its context decodes one batch segment after another until the queue stops
    void work(gi_ctx* g) {
        const SegmentRoom room{BATCH_MAX_STREAMS, g->P.seg_in_max, g->P.seg_out_max};
        std::vector<JobQueue::Item> items;
        std::vector<gi_job*> js;
        while (queue.take(room, items)) {
            const Stopwatch t0;
            js.resize(items.size());
            for (size_t i = 0; i < items.size(); i++) {
                js[i] = items[i].job;
                js[i]->queued_s = std::chrono::duration<double>(t0.origin() - items[i].queued).count();
            }
            if (js.size() == 1 && JobSizes::alone(js[0], room))
                decode_alone(g, js[0]);
            else
                run_batch(g, js.data(), js.size());
            const double sec = t0.seconds();
            for (gi_job* j : js) j->seconds = sec;
            queue.finish(items);
        }
    }
    void shutdown() {
        queue.stop();
        for (auto& t : threads)
            if (t.joinable()) t.join();
        threads.clear();
        for (gi_ctx* g : ctxs) gi_destroy(g);
        ctxs.clear();
    }
};

extern "C" {

// At most four workers: each context runs two HIP streams and a process gets four hardware
// queues (GPU_MAX_HW_QUEUES' default), so beyond two contexts streams share queues, and a
// finder graph queued behind another context's decode waits for it; more would only add waits.
constexpr int MAX_BATCH_WORKERS = 4;

GI_API int gi_batch_create(const gi_config* cfg, int workers, gi_batch** out, char* err, size_t err_len) {
    if (out == nullptr) return GI_ERR_ARG;
    *out = nullptr;
    if (workers < 1 || workers > MAX_BATCH_WORKERS) {
        copy_message(err, err_len, "workers must be 1.." + std::to_string(MAX_BATCH_WORKERS));
        return GI_ERR_ARG;
    }
    auto b = std::make_unique<gi_batch>();
    for (int i = 0; i < workers; i++) {
        gi_ctx* g = nullptr;
        const int rc = gi_create(cfg, &g, err, err_len);
        if (rc != GI_OK) {
            b->shutdown();
            return rc;
        }
        b->ctxs.push_back(g);
    }
    gi_batch* raw = b.get();
    for (gi_ctx* g : b->ctxs) b->threads.emplace_back([raw, g] { raw->work(g); });
    *out = b.release();
    return GI_OK;
}

GI_API int gi_batch_run(gi_batch* b, gi_job* jobs, uint64_t n) {
    if (b == nullptr || (jobs == nullptr && n)) return GI_ERR_ARG;
    b->queue.run(jobs, n, [](const gi_job* j) { return j->in_size; });
    for (uint64_t i = 0; i < n; i++)
        if (jobs[i].status != GI_OK) return jobs[i].status;
    return GI_OK;
}

GI_API int gi_batch_submit(gi_batch* b, gi_job* job) {
    if (b == nullptr || job == nullptr) return GI_ERR_ARG;
    b->queue.submit(job);
    return GI_OK;
}

GI_API gi_job* gi_batch_wait(gi_batch* b, double timeout_s) { return b ? b->queue.wait_done(timeout_s) : nullptr; }

GI_API int gi_batch_workers(const gi_batch* b) { return b ? (int)b->ctxs.size() : 0; }

GI_API int gi_batch_kernel_times(gi_batch* b, gi_kernel_time* out, int max, uint64_t* segments, uint64_t* alone,
                                 int reset) {
    if (b == nullptr) return 0;
    uint64_t launches[KID_COUNT] = {}, seg = 0, al = 0;
    double seconds[KID_COUNT] = {};
    for (gi_ctx* g : b->ctxs) {
        std::lock_guard<std::mutex> lk(g->acc_mu);
        for (int k = 0; k < KID_COUNT; k++) {
            launches[k] += g->acc_launches[k];
            seconds[k] += g->acc_seconds[k];
            if (reset) g->acc_launches[k] = 0, g->acc_seconds[k] = 0;
        }
        seg += g->acc_segments;
        al += g->acc_alone;
        if (reset) g->acc_segments = g->acc_alone = 0;
    }
    if (segments) *segments = seg;
    if (alone) *alone = al;
    int n = 0;
    for (int k = 0; k < KID_COUNT && out && n < max; k++) {
        if (!launches[k]) continue;
        out[n].name = KERNEL_NAMES[k];
        out[n].launches = launches[k];
        out[n].seconds = seconds[k];
        n++;
    }
    return n;
}

GI_API void gi_batch_destroy(gi_batch* b) {
    if (!b) return;
    b->shutdown();
    delete b;
}

}  // extern "C"

// ------------------------------------------------------------------ serving

namespace {

// Device memory of one allocation, handed out first fit in 256-byte units and given back
// (neighbouring free stretches merged). Not thread-safe: the server's lock guards it.
class DeviceArena {
  public:
    ~DeviceArena() {
        if (base_) (void)hipFree(base_);
    }
    void init(uint64_t bytes) {
        HIP_OK(hipMalloc(&base_, bytes));
        size_ = bytes;
        free_[0] = bytes;
    }
    static uint64_t units(uint64_t bytes) { return (std::max<uint64_t>(bytes, 1) + 255) / 256 * 256; }
    // the offset of `bytes` (whole units), or UINT64_MAX when no free stretch holds them
    uint64_t take(uint64_t bytes) {
        bytes = units(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const uint64_t off = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return UINT64_MAX;
    }
    void give(uint64_t off, uint64_t bytes) {
        bytes = units(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint8_t* at(uint64_t off) const { return (uint8_t*)base_ + off; }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }

  private:
    void* base_ = nullptr;
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> bytes
};

void set_error(gi_request* r, const std::string& why) { snprintf(r->error, sizeof r->error, "%s", why.c_str()); }

}  // namespace

struct gi_server {
    using Clock = std::chrono::steady_clock;
    // an admitted request: its device memory and what its trailer says
    struct Flight {
        gi_request* r;
        gi_job job;
        uint64_t in_off, in_bytes, out_off, out_bytes;
        uint32_t crc = 0, isize = 0;
        Clock::time_point submitted;
    };
    gi_batch* batch = nullptr;
    int device = 0;
    // uploads and downloads on streams of their own: on one stream, an upload waited for the
    // downloads queued before it (and the other way round), and a large output held up every
    // admission behind it
    hipStream_t st_up = nullptr, st_down = nullptr;
    DeviceArena arena;
    mutable std::mutex mu;
    std::condition_variable cv_work, cv_mem, cv_done;
    std::deque<std::pair<gi_request*, Clock::time_point>> waiting;  // submitted, not admitted
    std::deque<gi_request*> done;
    size_t flying = 0;
    bool stopping = false;
    std::thread admitter, completer;

    static double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

    void finish(gi_request* r, int status, Clock::time_point submitted) {
        r->status = status;
        r->done_s = since(submitted);
        std::lock_guard<std::mutex> lk(mu);
        done.push_back(r);
        cv_done.notify_all();
    }

    // admission, in arrival order: the head request waits until its device memory is free
    void admit_loop() {
        for (;;) {
            gi_request* r;
            Clock::time_point t0;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !waiting.empty(); });
                if (stopping) return;
                r = waiting.front().first;
                t0 = waiting.front().second;
            }
            // the Deflate data (a gzip member's body) and the output room
            uint64_t body = 0, body_size = r->in_size, room = r->out_cap;
            uint32_t crc = 0, isize = 0;
            int bad = GI_OK;
            if (r->gzip) {
                gzf::MemberHeader h;
                std::string why;
                if (!gzf::read_header(r->in, r->in_size, h, why)) {
                    set_error(r, why);
                    bad = GI_ERR_DATA;
                } else if (r->in_size < h.len + 8) {
                    set_error(r, "the member ends before its trailer");
                    bad = GI_ERR_DATA;
                } else {
                    body = h.len;
                    body_size = r->in_size - h.len - 8;
                    crc = gzf::le32(r->in + r->in_size - 8);
                    isize = gzf::le32(r->in + r->in_size - 4);
                    room = std::min<uint64_t>(room, isize);
                }
            }
            const uint64_t need_in = body_size + 64, need_out = room;
            if (bad == GI_OK && DeviceArena::units(need_in) + DeviceArena::units(need_out) > arena.size()) {
                set_error(r, "the request needs more device memory than the server's whole budget");
                bad = GI_ERR_ARG;
            }
            if (bad != GI_OK) {
                {
                    std::lock_guard<std::mutex> lk(mu);
                    waiting.pop_front();
                }
                finish(r, bad, t0);
                continue;
            }
            auto f = std::make_unique<Flight>();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_mem.wait(lk, [&] {
                    if (stopping) return true;
                    f->in_off = arena.take(need_in);
                    if (f->in_off == UINT64_MAX) return false;
                    f->out_off = arena.take(need_out);
                    if (f->out_off != UINT64_MAX) return true;
                    arena.give(f->in_off, need_in);
                    return false;
                });
                if (stopping) return;
                waiting.pop_front();
                flying++;
            }
            r->admitted_s = since(t0);
            f->r = r;
            f->in_bytes = need_in;
            f->out_bytes = need_out;
            f->crc = crc;
            f->isize = isize;
            f->submitted = t0;
            int rc = GI_OK;
            try {
                HIP_OK(hipSetDevice(device));
                if (body_size)
                    HIP_OK(hipMemcpyAsync(arena.at(f->in_off), r->in + body, body_size, hipMemcpyHostToDevice, st_up));
                HIP_OK(hipStreamSynchronize(st_up));
            } catch (const GpuFault& e) {
                set_error(r, e.what());
                rc = GI_ERR_GPU;
            }
            if (rc != GI_OK) {
                release(*f);
                finish(r, rc, t0);
                continue;
            }
            gi_job& j = f->job;
            memset(&j, 0, sizeof j);
            j.d_in = arena.at(f->in_off);
            j.in_size = body_size;
            j.d_out = arena.at(f->out_off);
            j.out_cap = room;
            j.out_hint = room;
            j.user = f.get();
            gi_batch_submit(batch, &f.release()->job);
        }
    }

    void release(Flight& f) {
        std::lock_guard<std::mutex> lk(mu);
        arena.give(f.in_off, f.in_bytes);
        arena.give(f.out_off, f.out_bytes);
        flying--;
        cv_mem.notify_all();
    }

    // completion: outputs to host memory, gzip trailers checked, device memory given back
    void complete_loop() {
        for (;;) {
            gi_job* j = gi_batch_wait(batch, 0.05);
            if (!j) {
                std::lock_guard<std::mutex> lk(mu);
                if (stopping && flying == 0) return;
                continue;
            }
            std::unique_ptr<Flight> f((Flight*)j->user);
            gi_request* r = f->r;
            int status = j->status;
            if (status == GI_OK) {
                try {
                    HIP_OK(hipSetDevice(device));
                    if (j->out_size)
                        HIP_OK(hipMemcpyAsync(r->out, j->d_out, j->out_size, hipMemcpyDeviceToHost, st_down));
                    HIP_OK(hipStreamSynchronize(st_down));
                    r->out_size = j->out_size;
                    r->crc32 = j->crc32;
                    if (r->gzip && ((uint32_t)j->out_size != f->isize || j->crc32 != f->crc)) {
                        set_error(r, "the output does not match the member's trailer");
                        status = GI_ERR_DATA;
                    }
                } catch (const GpuFault& e) {
                    set_error(r, e.what());
                    status = GI_ERR_GPU;
                }
            } else {
                set_error(r, status == GI_ERR_ARG ? "the output is larger than its room" : "the stream did not decode");
            }
            release(*f);
            finish(r, status, f->submitted);
        }
    }
};

extern "C" {

GI_API int gi_serve_create(const gi_config* cfg, int workers, uint64_t hbm_budget, gi_server** out, char* err,
                           size_t err_len) {
    if (out == nullptr || cfg == nullptr) return GI_ERR_ARG;
    *out = nullptr;
    if (hbm_budget < (1u << 20)) {
        copy_message(err, err_len, "the HBM budget must be at least 1 MiB");
        return GI_ERR_ARG;
    }
    auto s = std::make_unique<gi_server>();
    const int rc = gi_batch_create(cfg, workers, &s->batch, err, err_len);
    if (rc != GI_OK) return rc;
    s->device = cfg->size >= 8 ? cfg->device : 0;
    try {
        HIP_OK(hipSetDevice(s->device));
        HIP_OK(hipStreamCreateWithFlags(&s->st_up, hipStreamNonBlocking));
        HIP_OK(hipStreamCreateWithFlags(&s->st_down, hipStreamNonBlocking));
        s->arena.init(hbm_budget);
    } catch (const GpuFault& e) {
        copy_message(err, err_len, e.what());
        if (s->st_up) (void)hipStreamDestroy(s->st_up);
        if (s->st_down) (void)hipStreamDestroy(s->st_down);
        gi_batch_destroy(s->batch);
        return GI_ERR_GPU;
    }
    gi_server* raw = s.get();
    s->admitter = std::thread([raw] { raw->admit_loop(); });
    s->completer = std::thread([raw] { raw->complete_loop(); });
    *out = s.release();
    return GI_OK;
}

GI_API int gi_serve_submit(gi_server* s, gi_request* r) {
    if (s == nullptr || r == nullptr || (r->in == nullptr && r->in_size) || (r->out == nullptr && r->out_cap))
        return GI_ERR_ARG;
    r->out_size = 0;
    r->crc32 = 0;
    r->status = GI_OK;
    r->admitted_s = r->done_s = 0;
    r->error[0] = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    s->waiting.push_back({r, gi_server::Clock::now()});
    s->cv_work.notify_all();
    return GI_OK;
}

GI_API gi_request* gi_serve_wait(gi_server* s, double timeout_s) {
    if (s == nullptr) return nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(std::max(0.0, timeout_s)), [&] { return !s->done.empty(); }))
        return nullptr;
    gi_request* r = s->done.front();
    s->done.pop_front();
    return r;
}

GI_API uint64_t gi_serve_in_use(const gi_server* s, uint64_t* peak) {
    if (s == nullptr) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    if (peak) *peak = s->arena.peak();
    return s->arena.used();
}

GI_API void gi_serve_destroy(gi_server* s) {
    if (s == nullptr) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv_work.notify_all();
        s->cv_mem.notify_all();
    }
    if (s->admitter.joinable()) s->admitter.join();
    if (s->completer.joinable()) s->completer.join();
    gi_batch_destroy(s->batch);
    if (s->st_up) (void)hipStreamDestroy(s->st_up);
    if (s->st_down) (void)hipStreamDestroy(s->st_down);
    delete s;
}

GI_API int gi_inflate_many(gi_ctx* g, const uint8_t* in, gi_stream* streams, uint64_t n, uint8_t* out, uint64_t out_cap,
                           gi_stats* stats) {
    if (g == nullptr || (n && (streams == nullptr || in == nullptr))) return GI_ERR_ARG;
    g->err.clear();
    if (stats) memset(stats, 0, sizeof *stats);
    for (uint64_t i = 0; i < n; i++) {
        gi_stream& s = streams[i];
        s.out_size = s.in_used = 0;
        s.crc32 = 0;
        s.status = GI_OK;
        if (s.out_offset > out_cap || s.out_room > out_cap - s.out_offset || (s.out_room && !out)) {
            g->err = "stream " + std::to_string(i) + "'s output lies outside the output buffer";
            return GI_ERR_ARG;
        }
    }
    const Stopwatch clock;
    const SegmentRoom room{BATCH_MAX_STREAMS, g->P.seg_in_max, g->P.seg_out_max};
    auto in_bytes = [](const gi_stream& s) { return (s.size + 15) / 16 * 16 + 16; };
    auto out_bytes = [](const gi_stream& s) { return (s.out_room + BATCH_ALIGN - 1) / BATCH_ALIGN * BATCH_ALIGN; };
    uint32_t groups = 0;
    for (uint64_t i = 0; i < n;) {
        // a group: the streams from i on while they fit one batch segment; one over a quarter of
        // a segment's input or output goes alone
        uint64_t j = i, ib = 0, ob = 0;
        bool alone = false;
        while (j < n && j - i < BATCH_MAX_STREAMS) {
            const gi_stream& s = streams[j];
            if (s.size > room.in_bytes / 4 || s.out_room > room.out_bytes / 4) {
                if (j == i) {
                    alone = true;
                    j++;
                }
                break;
            }
            if (ib + in_bytes(s) > room.in_bytes || ob + out_bytes(s) > room.out_bytes) break;
            ib += in_bytes(s);
            ob += out_bytes(s);
            j++;
        }
        const uint64_t m = j - i;
        groups++;
        // their inputs packed into device memory (16-byte aligned, zero padded), room for their outputs
        std::vector<uint64_t> at_in(m), at_out(m);
        uint64_t tin = 0, tout = 0;
        for (uint64_t k = 0; k < m; k++) {
            at_in[k] = tin;
            at_out[k] = tout;
            tin += in_bytes(streams[i + k]);
            tout += out_bytes(streams[i + k]);
        }
        void* d_in = nullptr;
        void* d_out = nullptr;
        std::vector<gi_job> jobs(m);
        int rc = on_context(*g, [&] {
            HIP_OK(hipSetDevice(g->device));
            HIP_OK(hipMalloc(&d_in, tin + 64));
            HIP_OK(hipMalloc(&d_out, std::max<uint64_t>(tout, 16)));
            HIP_OK(hipMemsetAsync(d_in, 0, tin + 64, g->st));
            for (uint64_t k = 0; k < m; k++) {
                const gi_stream& s = streams[i + k];
                if (s.size)
                    HIP_OK(hipMemcpyAsync((uint8_t*)d_in + at_in[k], in + s.offset, s.size, hipMemcpyHostToDevice, g->st));
            }
            HIP_OK(hipStreamSynchronize(g->st));
            return GI_OK;
        });
        if (rc == GI_OK) {
            std::vector<gi_job*> ptrs(m);
            for (uint64_t k = 0; k < m; k++) {
                const gi_stream& s = streams[i + k];
                gi_job& jb = jobs[k];
                memset(&jb, 0, sizeof jb);
                jb.d_in = (const uint8_t*)d_in + at_in[k];
                jb.in_size = s.size;
                jb.d_out = (uint8_t*)d_out + at_out[k];
                jb.out_cap = s.out_room;
                jb.out_hint = s.out_room;
                ptrs[k] = &jb;
            }
            if (alone)
                decode_alone(g, ptrs[0]);
            else
                run_batch(g, ptrs.data(), m);
            rc = on_context(*g, [&] {
                // the outputs in one copy, then to their places
                std::vector<uint8_t> host(tout);
                if (tout) HIP_OK(hipMemcpy(host.data(), d_out, tout, hipMemcpyDeviceToHost));
                for (uint64_t k = 0; k < m; k++) {
                    const gi_job& jb = jobs[k];
                    gi_stream& s = streams[i + k];
                    s.status = jb.status;
                    if (jb.status != GI_OK) continue;
                    s.out_size = jb.out_size;
                    s.in_used = jb.in_used;
                    s.crc32 = jb.crc32;
                    if (jb.out_size) memcpy(out + s.out_offset, host.data() + at_out[k], jb.out_size);
                }
                return GI_OK;
            });
        }
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        if (rc != GI_OK) {
            for (uint64_t k = i; k < n; k++) streams[k].status = rc;
            break;
        }
        i = j;
    }
    int first = GI_OK;
    for (uint64_t i = 0; i < n; i++) {
        if (stats && streams[i].status == GI_OK) {
            stats->in_used += streams[i].in_used;
            stats->out_size += streams[i].out_size;
        }
        if (first == GI_OK && streams[i].status != GI_OK) {
            first = streams[i].status;
            if (g->err.empty()) g->err = "stream " + std::to_string(i) + " did not decode";
        }
    }
    if (stats) {
        stats->segments = groups;
        stats->jobs = n;
        stats->seconds = clock.seconds();
    }
    return first;
}

}  // extern "C"
