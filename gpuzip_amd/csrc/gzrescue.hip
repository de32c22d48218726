// gzrescue.hip - libgzrescue.so: the rescue of damaged gzip files and raw Deflate streams behind
// the reference's gr_* interface, ABI 4 (include/gzrescue.h; reference gpuinflate/rescue.h:38-107,
// gpuinflate/rescue.cpp:2767-2790), with its work on an MI355X (gfx950).
//
// The decisions (rescue_core.h: the pass walk, the checks after it, the inference, the emitter)
// run on the host, over this backend:
//   find            every dynamic block header the decoder would accept and that decodes on
//                   (libgpuinflate's gi_rescue_headers: the block finder, then a verifying decode
//                   from every plausible header, all on the GPU);
//   decode_headers  a decode from every header to where it lands on the next one: one GPU lane a
//                   header (k_rescue_decode), running the sequential decoder of seqdec.h, whose
//                   results are the reference decoder's field for field;
//   decode_other    the few decodes from elsewhere (a stream's start, a resync), on the host with
//                   the same decoder, their output then copied to the GPU;
//   resolve         each output byte's provenance (its value, or the lost-window variable it
//                   copies, or nothing) by pointer jumping over the whole output in HBM
//                   (k_prov_init, k_prov_jump, k_prov_split); it stays there, and what the
//                   decisions read off it is made there: per piece, exactness and the first
//                   variable at a multiple of 64 (k_sum_pieces); ranges of variables in the kept
//                   pieces (k_var_edges); each member's CRC-32 (k_crc_chunks, combined on the
//                   host); the bytes that copy variables (k_var_places: their tags to the
//                   host, and of those it picks, the places with their neighbours, k_pick_places);
//   emit            the output and its mask (k_emit), to the host through pinned buffers.
//
// A decode writes one 16-bit symbol per output byte: v <= 255 a literal, v > 255 the byte (v - 255)
// back. The reference decodes a piece in calls of at most SCRATCH symbols and resumes after each
// full buffer; a stored block that does not fit is counted again by the call that resumes, so the
// split points show in the results. A lane here is given less room than that at first: when it
// fills up before the reference's call would have, its call is run again from the same start with
// more room (four times as much, up to the reference's), so every call ends where the reference's
// does.
#define GR_BUILD
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpuinflate.h"
#include "../../include/gzrescue.h"
#include "rescue_core.h"

using gz::Start;

namespace {

struct GpuFault : std::runtime_error {
    using std::runtime_error::runtime_error;
};
inline void hip_check(hipError_t rc, const char* call) {
    if (rc != hipSuccess) throw GpuFault(std::string(call) + ": " + hipGetErrorString(rc));
}
#define HIP_OK(call) hip_check((call), #call)

// A device allocation, freed with its owner
class DevMem {
  public:
    DevMem() = default;
    explicit DevMem(size_t bytes) { reset(bytes); }
    ~DevMem() { release(); }
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    void reset(size_t bytes) {
        release();
        if (bytes) HIP_OK(hipMalloc(&p_, bytes));
        n_ = bytes;
    }
    // at least `bytes`, keeping nothing
    void reserve(size_t bytes) {
        if (bytes > n_) reset(bytes);
    }
    template <class T = void>
    T* get() const {
        return (T*)p_;
    }
    size_t size() const { return n_; }

  private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void* p_ = nullptr;
    size_t n_ = 0;
};

// ------------------------------------------------------------------ kernels

// One call of a decode: the reference decoder's run() from s, landing on targets t0.., writing at
// most cap symbols at out (an offset in the round's pool)
struct LaneJob {
    Start s;
    uint64_t out;
    uint32_t cap;
    int32_t t0;
};

// One lane a call. The decoder's state (its code tables, about 4 KB) lives in HBM, one slot a
// lane, zeroed before the launch; the lanes of a wavefront decode unrelated blocks, so they share
// no table.
__global__ __launch_bounds__(64) void k_rescue_decode(const uint8_t* in, uint64_t size, const Start* tg, int ntg,
                                                      const LaneJob* jobs, int n, uint16_t* pool,
                                                      seqd::Inflater* work, seqd::Outcome* res,
                                                      gz::StoredList stored) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const LaneJob j = jobs[i];
    const int t0 = j.t0 < ntg ? j.t0 : ntg;
    res[i] = work[i].run(in, size, j.s, tg + t0, ntg - t0, pool + j.out, j.cap, false, UINT64_MAX, &stored, j.out);
}

// The stored blocks the decoders listed: a workgroup a block, 8 bytes a thread a step read as one
// load and written as one 16-byte store of 8 symbols, the tail byte by byte
__global__ __launch_bounds__(256) void k_rescue_stored(const gz::StoredCopy* list, const unsigned* count, unsigned cap,
                                                       const uint8_t* in, uint16_t* pool, unsigned long long* copied) {
    typedef uint2 __attribute__((aligned(1))) Bytes8;
    typedef uint4 __attribute__((aligned(2))) Syms8;
    const unsigned n = min(*count, cap);
    for (unsigned b = blockIdx.x; b < n; b += gridDim.x) {
        const gz::StoredCopy c = list[b];
        if (threadIdx.x == 0) atomicAdd(copied, (unsigned long long)c.len);
        const uint32_t whole = c.len & ~7u;
        for (uint32_t k = 8 * threadIdx.x; k < whole; k += 8 * blockDim.x) {
            const Bytes8 v = *(const Bytes8*)(in + c.src + k);
            Syms8 q;
            q.x = (v.x & 0xFFu) | (v.x & 0xFF00u) << 8;
            q.y = (v.x >> 16 & 0xFFu) | (v.x >> 8 & 0xFF0000u);
            q.z = (v.y & 0xFFu) | (v.y & 0xFF00u) << 8;
            q.w = (v.y >> 16 & 0xFFu) | (v.y >> 8 & 0xFF0000u);
            *(Syms8*)(pool + c.dst + k) = q;
        }
        for (uint32_t k = whole + threadIdx.x; k < c.len; k += blockDim.x) pool[c.dst + k] = in[c.src + k];
    }
}

// Provenance entries, 64 bits an output byte. Settled (bit 63 clear): the byte in bits 32..39 and
// its tag in bits 0..31 (EXACT, a variable, or 0 for nothing known); zero is a settled unknown
// byte, so a cleared array holds the bytes no stretch covers. Pending (bit 63 set): the byte is
// that of the output byte at bits 0..61, and bit 62 says that its tag becomes 0 (a blank tail
// lies on the way).
constexpr uint64_t PENDING = 1ull << 63;
constexpr uint64_t BLANKED = 1ull << 62;
constexpr uint64_t TARGET = BLANKED - 1;
constexpr uint32_t PROV_CHUNK = 8192;  // output bytes of a stretch one workgroup sets up
constexpr int JUMP_HOPS = 8;           // links a lane follows in one launch

struct DevSpan {
    const uint16_t* src;  // the decode's symbols
    uint64_t out, len;
    uint64_t window;      // bytes of the output before `out` its references may reach (has_window)
    uint64_t tail;        // the first offset of its blank tail (len: none)
    uint32_t var_base;
    uint32_t has_window;
};
struct SpanChunk {
    uint32_t span;
    uint32_t pad;
    uint64_t first;
};

__global__ __launch_bounds__(256) void k_prov_init(const DevSpan* spans, const SpanChunk* chunks, uint64_t* ent) {
    const SpanChunk c = chunks[blockIdx.x];
    const DevSpan sp = spans[c.span];
    const uint64_t end = min(sp.len, c.first + PROV_CHUNK);
    for (uint64_t i = c.first + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t v = sp.src[i];
        const uint64_t x = sp.out + i;
        const bool blank = i >= sp.tail;
        uint64_t e;
        if (v <= gz::REF) {
            e = (uint64_t)v << 32 | (blank ? 0u : rsc::Provenance::EXACT);
        } else {
            const uint64_t d = v - gz::REF;
            if (d <= i || (sp.has_window && d - i <= sp.window)) {
                e = PENDING | (blank ? BLANKED : 0) | (x - d);
            } else {
                const uint32_t tag = sp.var_base && d - i <= gz::WINDOW ? sp.var_base + (uint32_t)(d - i) : 0u;
                e = blank ? 0 : tag;
            }
        }
        ent[x] = e;
    }
}

// Every pending entry follows up to JUMP_HOPS links, taking over what it finds: a settled entry
// settles it (tag 0 if a blank tail was crossed), a pending one moves it further back. Entries
// are read and written in place by many lanes: any value an entry ever holds describes the same
// byte, so a lane that reads an older one only gets less far. *left counts what stays pending.
__global__ __launch_bounds__(256) void k_prov_jump(uint64_t* ent, uint64_t total, unsigned long long* left) {
    unsigned long long pending = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
        uint64_t e = __hip_atomic_load(ent + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(e & PENDING)) continue;
        for (int h = 0; h < JUMP_HOPS && (e & PENDING); h++) {
            const uint64_t blank = e & BLANKED;
            const uint64_t f = __hip_atomic_load(ent + (e & TARGET), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            e = (f & PENDING) ? f | blank : blank ? f & 0xFFFFFFFF00000000ull : f;
        }
        __hip_atomic_store(ent + x, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pending += (e & PENDING) ? 1 : 0;
    }
    // one atomic a wavefront
    for (int o = 32; o > 0; o >>= 1) pending += __shfl_down(pending, o, 64);
    if ((threadIdx.x & 63) == 0 && pending) atomicAdd(left, pending);
}

__global__ __launch_bounds__(256) void k_prov_split(const uint64_t* ent, uint64_t total, uint8_t* byte, uint32_t* tag) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
        const uint64_t e = ent[x];
        byte[x] = (uint8_t)(e >> 32);
        tag[x] = (uint32_t)e;
    }
}

// ------------------------------------------------------------------ what the decisions read off a pass

// A stretch of the provenance [at, at + len), and a tile of one: stretch idx, its offsets
// [first, first + tile size)
struct Stretch {
    uint64_t at, len;
};
struct Tile {
    uint32_t idx;
    uint32_t pad;
    uint64_t first;
};
constexpr uint32_t SUM_TILE = 65536;   // bytes a workgroup summarizes
constexpr uint32_t SCAN_TILE = 4096;   // bytes a wavefront compacts, in order
constexpr uint32_t CRC_CHUNK = 65536;  // bytes a wavefront's CRC covers: 1 KiB a lane
constexpr uint32_t EMIT_TILE = 65536;  // output bytes a workgroup emits

__device__ __forceinline__ bool is_var(uint32_t t) { return t != rsc::Provenance::EXACT && t != 0; }

// Per piece: whether any byte is not EXACT, and the first offset, a multiple of 64, whose byte
// copies a variable (all ones: none); one atomic a wavefront
__global__ __launch_bounds__(256) void k_sum_pieces(const uint32_t* tag, const Stretch* pieces, const Tile* tiles,
                                                    unsigned* inexact, unsigned long long* var64) {
    const Tile c = tiles[blockIdx.x];
    const Stretch p = pieces[c.idx];
    const uint64_t end = min(p.len, c.first + SUM_TILE);
    bool bad = false;
    unsigned long long first = ~0ull;
    for (uint64_t i = c.first + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t t = tag[p.at + i];
        bad |= t != rsc::Provenance::EXACT;
        if ((i & 63) == 0 && is_var(t) && i < first) first = i;
    }
    const bool any_bad = __ballot(bad) != 0;
    for (int o = 32; o > 0; o >>= 1) first = min(first, (unsigned long long)__shfl_down(first, o, 64));
    if ((threadIdx.x & 63) == 0) {
        if (any_bad) atomicOr(inexact + c.idx, 1u);
        if (first != ~0ull) atomicMin(var64 + c.idx, first);
    }
}

// The ranges of bytes that copy variables, in kept pieces: a range begins at a variable byte whose
// predecessor (in the piece) is not one, ends after one whose successor is not. One wavefront a
// tile counts both (WRITE false), or writes them at the tile's offsets in order (WRITE true).
template <bool WRITE>
__global__ __launch_bounds__(64) void k_var_edges(const uint32_t* tag, const Stretch* pieces, const Tile* tiles,
                                                  unsigned* n_starts, unsigned* n_ends, const unsigned* at_starts,
                                                  const unsigned* at_ends, uint64_t* starts, uint64_t* ends) {
    const Tile c = tiles[blockIdx.x];
    const Stretch p = pieces[c.idx];
    const uint64_t end = min(p.len, c.first + SCAN_TILE);
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1;
    unsigned ns = WRITE ? at_starts[blockIdx.x] : 0, ne = WRITE ? at_ends[blockIdx.x] : 0;
    for (uint64_t base = c.first; base < end; base += 64) {
        const uint64_t i = base + lane;
        bool st = false, en = false;
        if (i < end) {
            const bool v = is_var(tag[p.at + i]);
            st = v && !(i > 0 && is_var(tag[p.at + i - 1]));
            en = v && !(i + 1 < p.len && is_var(tag[p.at + i + 1]));
        }
        const uint64_t ms = __ballot(st), me = __ballot(en);
        if (WRITE) {
            if (st) starts[ns + __popcll(ms & below)] = i;
            if (en) ends[ne + __popcll(me & below)] = i + 1;
        }
        ns += __popcll(ms);
        ne += __popcll(me);
    }
    if (!WRITE && lane == 0) {
        n_starts[blockIdx.x] = ns;
        n_ends[blockIdx.x] = ne;
    }
}

// Every byte of the stretches that copies a variable, with its cells (the bytes around it in its
// stretch: a byte, VAR + a variable, or NONE): counted per tile, or written in order
template <bool WRITE>
__global__ __launch_bounds__(64) void k_var_places(const uint8_t* byte, const uint32_t* tag, const Stretch* st,
                                                   const Tile* tiles, unsigned* counts, const unsigned* at,
                                                   rsc::Place* places, uint32_t* tags) {
    const Tile c = tiles[blockIdx.x];
    const Stretch s = st[c.idx];
    const uint64_t end = min(s.len, c.first + SCAN_TILE);
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1;
    unsigned n = WRITE ? at[blockIdx.x] : 0;
    for (uint64_t base = c.first; base < end; base += 64) {
        const uint64_t x = base + lane;
        const bool v = x < end && is_var(tag[s.at + x]);
        const uint64_t m = __ballot(v);
        if (WRITE && v) {
            const unsigned k = n + __popcll(m & below);
            rsc::Place& pl = places[k];
            pl.pos = s.at + x;
            pl.tag = tag[s.at + x];
            tags[k] = pl.tag;
            auto cell = [&](uint64_t j) -> uint32_t {
                const uint32_t t = tag[s.at + j];
                return t == rsc::Provenance::EXACT ? byte[s.at + j] : t ? rsc::Occ::VAR + t : rsc::Occ::NONE;
            };
            for (int q = 1; q <= rsc::Occ::BEFORE; q++) pl.cell[q - 1] = x >= (uint64_t)q ? cell(x - q) : rsc::Occ::NONE;
            for (int q = 1; q <= rsc::Occ::AFTER; q++)
                pl.cell[rsc::Occ::BEFORE + q - 1] = x + q < s.len ? cell(x + q) : rsc::Occ::NONE;
        }
        n += __popcll(m);
    }
    if (!WRITE && lane == 0) counts[blockIdx.x] = n;
}

// the places picked out of all (indices in order)
__global__ __launch_bounds__(256) void k_pick_places(const rsc::Place* all, const uint32_t* picked, unsigned n,
                                                     rsc::Place* out) {
    for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) out[j] = all[picked[j]];
}

// a(x) b(x) modulo the CRC polynomial, reflected (bit 31: x^0)
__device__ __forceinline__ uint32_t mulmodp(uint32_t a, uint32_t b) {
    uint32_t p = 0;
    for (uint32_t m = 1u << 31; m; m >>= 1) {
        if (a & m) p ^= b;
        b = (b & 1) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
    }
    return p;
}
struct CrcShifts {
    uint32_t k[64];  // x^(8 * 1024 * j) modulo P: lane j's CRC moved past the 1 KiB pieces after it
};

// The CRC-32 of each 64 KiB chunk of bytes[base, base + 65536 n): each lane the CRC of its KiB (a
// byte table in LDS), moved past the lanes after it and summed over the wavefront
__global__ __launch_bounds__(64) void k_crc_chunks(const uint8_t* bytes, uint64_t base, uint64_t n, CrcShifts sh,
                                                   uint32_t* out) {
    __shared__ uint32_t table[256];
    for (uint32_t t = threadIdx.x; t < 256; t += blockDim.x) {
        uint32_t c = t;
        for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
        table[t] = c;
    }
    __syncthreads();
    typedef uint32_t __attribute__((aligned(1))) Word;
    for (uint64_t ch = blockIdx.x; ch < n; ch += gridDim.x) {
        const uint8_t* p = bytes + base + ch * CRC_CHUNK + threadIdx.x * 1024u;
        uint32_t c = 0xFFFFFFFFu;
        for (int i = 0; i < 1024; i += 4) {
            uint32_t w = *(const Word*)(p + i);
            for (int k = 0; k < 4; k++, w >>= 8) c = table[(c ^ w) & 255] ^ (c >> 8);
        }
        c = mulmodp(~c, sh.k[63 - threadIdx.x]);
        for (int o = 32; o > 0; o >>= 1) c ^= __shfl_xor(c, o, 64);
        if (threadIdx.x == 0) out[ch] = c;
    }
}

// A piece to emit: the provenance's bytes [at, at + min(decoded, length)), then fill to length;
// `off` its first output byte
struct EPiece {
    uint64_t at, decoded, length, off;
    uint32_t suspect, pad;
};
struct ECounts {
    unsigned long long known, exact, suspect, unknown;
};

// The output bytes and mask of the tiles [t0, t0 + gridDim.x) (EMIT_TILE output bytes each) into
// out/mask from output byte w0 on; counts per piece. A variable's byte is its inferred value
// (binary search of vars[0, nv), sorted), or the fill byte when it has none.
__global__ __launch_bounds__(256) void k_emit(const uint8_t* byte, const uint32_t* tag, const EPiece* pieces, int np,
                                              uint64_t total, uint64_t t0, uint64_t w0, const uint32_t* vars,
                                              const uint8_t* vals, int nv, uint8_t fill, uint8_t* out, uint8_t* mask,
                                              ECounts* counts) {
    const uint64_t ts = (t0 + blockIdx.x) * EMIT_TILE, te = min(total, ts + EMIT_TILE);
    if (ts >= te) return;
    // the last piece that starts at or before ts
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pieces[mid].off <= ts) lo = mid;
        else hi = mid - 1;
    }
    for (int k = lo; k < np && pieces[k].off < te; k++) {
        const EPiece p = pieces[k];
        const uint64_t a = max(ts, p.off), b = min(te, p.off + p.length);
        if (a >= b) continue;
        const uint64_t e = min(p.decoded, p.length);
        unsigned long long known = 0, exact = 0, suspect = 0, unknown = 0;
        for (uint64_t o = a + threadIdx.x; o < b; o += blockDim.x) {
            const uint64_t i = o - p.off;
            uint8_t x = fill, m = GR_UNKNOWN;
            if (i < e) {
                const uint32_t t = tag[p.at + i];
                if (t == rsc::Provenance::EXACT) {
                    x = byte[p.at + i];
                    m = p.suspect ? GR_SUSPECT : GR_EXACT;
                } else if (t) {
                    int l = 0, h = nv;
                    while (l < h) {
                        const int mid = (l + h) >> 1;
                        if (vars[mid] < t) l = mid + 1;
                        else h = mid;
                    }
                    if (l < nv && vars[l] == t) {
                        x = vals[l];
                        m = p.suspect ? GR_SUSPECT : GR_INFERRED;
                    }
                }
            }
            known += m != GR_UNKNOWN;
            exact += m == GR_EXACT;
            suspect += m == GR_SUSPECT;
            unknown += m == GR_UNKNOWN;
            out[o - w0] = x;
            mask[o - w0] = m;
        }
        for (int s = 32; s > 0; s >>= 1) {
            known += __shfl_down(known, s, 64);
            exact += __shfl_down(exact, s, 64);
            suspect += __shfl_down(suspect, s, 64);
            unknown += __shfl_down(unknown, s, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            if (known) atomicAdd(&counts[k].known, known);
            if (exact) atomicAdd(&counts[k].exact, exact);
            if (suspect) atomicAdd(&counts[k].suspect, suspect);
            if (unknown) atomicAdd(&counts[k].unknown, unknown);
        }
    }
}

unsigned grid_for(uint64_t n, unsigned block, unsigned most = 16384) {
    const uint64_t g = (n + block - 1) / block;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, most));
}

// ------------------------------------------------------------------ the backend

// GPU time per kernel, from HIP events recorded around every launch on the rescue's stream
class KernelClock {
  public:
    explicit KernelClock(hipStream_t st) : st_(st) {}
    ~KernelClock() {
        for (Mark& m : open_) {
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
        }
    }
    template <class F>
    void time(const char* name, F&& launch) {
        Mark m{name, nullptr, nullptr};
        HIP_OK(hipEventCreate(&m.a));
        HIP_OK(hipEventCreate(&m.b));
        open_.push_back(m);
        HIP_OK(hipEventRecord(m.a, st_));
        launch();
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(m.b, st_));
    }
    // after the stream has been synchronized
    void settle() {
        for (Mark& m : open_) {
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, m.a, m.b));
            (void)hipEventDestroy(m.a);
            (void)hipEventDestroy(m.b);
            auto it = std::find_if(totals_.begin(), totals_.end(), [&](const gzr_kernel_time& k) { return !strcmp(k.name, m.name); });
            if (it == totals_.end()) it = totals_.insert(totals_.end(), gzr_kernel_time{m.name, 0, 0.0});
            it->launches++;
            it->seconds += ms * 1e-3;
        }
        open_.clear();
    }
    const std::vector<gzr_kernel_time>& totals() const { return totals_; }

  private:
    struct Mark {
        const char* name;
        hipEvent_t a, b;
    };
    hipStream_t st_;
    std::vector<Mark> open_;
    std::vector<gzr_kernel_time> totals_;
};

// the reference's call size: a call that fills this many symbols is resumed, as the reference's is
constexpr uint32_t CALL_CAP = rsc::SCRATCH + 2 * gz::MAX_SYMBOL_OUT;
constexpr uint32_t FIRST_CAP_MIN = 128u << 10;  // a stored block fits, whatever the call has written
constexpr uint32_t CAP_RATIO = 6;               // first room: symbols per compressed byte to the next header

// The process's GPU state: the header search's context and the stream the kernels run on. One
// rescue at a time uses it.
// Two pinned host buffers (output and mask each) that device-to-host copies take in turn
struct Staging {
    struct Buf {
        uint8_t* out = nullptr;
        uint8_t* mask = nullptr;
        hipEvent_t done = nullptr;
    } buf[2];
    size_t bytes = 0;
    void release() {
        for (Buf& b : buf) {
            if (b.out) (void)hipHostFree(b.out);
            if (b.mask) (void)hipHostFree(b.mask);
            if (b.done) (void)hipEventDestroy(b.done);
            b = Buf{};
        }
        bytes = 0;
    }
};

struct Device {
    std::mutex mu;
    gi_ctx* finder = nullptr;
    hipStream_t st = nullptr;
    int device = 0;
    Staging stage;

    Staging& staging(size_t bytes) {
        if (stage.bytes < bytes) {
            stage.release();
            for (Staging::Buf& b : stage.buf) {
                HIP_OK(hipHostMalloc((void**)&b.out, bytes, hipHostMallocDefault));
                HIP_OK(hipHostMalloc((void**)&b.mask, bytes, hipHostMallocDefault));
                HIP_OK(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
            }
            stage.bytes = bytes;
        }
        return stage;
    }

    void close() {
        stage.release();
        if (finder) gi_destroy(finder);
        finder = nullptr;
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
    }
    void open() {
        if (finder) return;
        HIP_OK(hipSetDevice(device));
        if (!st) HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        gi_config cfg{};
        cfg.size = sizeof cfg;
        cfg.device = device;
        cfg.segment_out_mb = 512;  // the search's segment: 1 GiB of input
        char err[256] = {0};
        if (gi_create(&cfg, &finder, err, sizeof err) != GI_OK) {
            finder = nullptr;
            throw GpuFault(std::string("gi_create: ") + err);
        }
    }
};

Device& the_device() {
    static Device d;
    return d;
}

class GpuRescue {
  public:
    explicit GpuRescue(Device& dev) : dev_(dev), clock_(dev.st) {}
    const std::vector<gzr_kernel_time>& kernel_times() const { return clock_.totals(); }
    // bytes the decoders left to the stored-block copy kernel
    uint64_t stored_bytes() const { return stored_bytes_; }

    void attach(const uint8_t* p, uint64_t n) {
        in_ = p;
        size_ = n;
        HIP_OK(hipSetDevice(dev_.device));
        // the decoders read whole 32-bit words: the padding the host buffer has (zeros) goes too
        d_in_.reset(n + 64);
        HIP_OK(hipMemcpyAsync(d_in_.get(), p, n + 64, hipMemcpyHostToDevice, dev_.st));
        // a stored block takes at least 5 input bytes; one a KiB is room for every stored block
        // of a file of them at full length (a longer list spills: lanes then copy themselves)
        stored_cap_ = (unsigned)std::min<uint64_t>(n / 1024 + 1024, 1u << 24);
        d_stored_.reset((size_t)stored_cap_ * sizeof(gz::StoredCopy));
        d_stored_count_.reset(sizeof(unsigned));
        d_copied_.reset(8);
        HIP_OK(hipStreamSynchronize(dev_.st));
    }

    void find(uint64_t from, std::vector<Start>& out) {
        out.clear();
        const uint64_t nbits = size_ * 8;
        if (from >= (nbits > 128 ? nbits - 128 : 0)) return;
        // the library searches up to 1 GiB of input at once: longer inputs in windows that overlap
        // by far more than a verifying decode reads (256 Ki symbols, at most 48 bits each)
        constexpr uint64_t WIN = 1ull << 30, MARGIN = 16ull << 20;
        uint64_t lo = from;  // bit: headers before it are taken
        for (uint64_t w0 = 0;;) {
            const uint64_t wn = std::min(WIN, size_ - w0);
            const bool last = w0 + wn == size_;
            const uint64_t hi = last ? UINT64_MAX : (w0 + wn - MARGIN) * 8;
            std::vector<uint64_t> got;
            search(in_ + w0, wn, lo - w0 * 8, got);
            for (uint64_t b : got)
                if (b + w0 * 8 >= lo && b + w0 * 8 < hi) out.push_back(rsc::at_block(b + w0 * 8));
            if (last) break;
            lo = hi;
            w0 = hi / 8 - 1;  // the window's first bit searched is 8 bits in, as `from` >= 1 asks
        }
    }

    void decode_headers(const std::vector<Start>& h, std::vector<rsc::Decoded>& res) {
        const size_t n = h.size();
        res.assign(n, rsc::Decoded{});
        const size_t first = outs_.size();
        outs_.resize(first + n);
        if (!n) return;
        HIP_OK(hipSetDevice(dev_.device));
        d_targets_.reset(n * sizeof(Start));
        HIP_OK(hipMemcpyAsync(d_targets_.get(), h.data(), n * sizeof(Start), hipMemcpyHostToDevice, dev_.st));
        struct Call {
            size_t id;
            Start s;
            uint32_t cap;
        };
        std::vector<Call> todo(n);
        for (size_t i = 0; i < n; i++) {
            const uint64_t next = i + 1 < n ? h[i + 1].bit : size_ * 8;
            const uint64_t bytes = next > h[i].bit ? (next - h[i].bit + 7) / 8 : 0;
            const uint64_t want = bytes * CAP_RATIO + FIRST_CAP_MIN;
            todo[i] = Call{i, h[i], (uint32_t)std::min<uint64_t>(want, CALL_CAP)};
        }
        std::vector<std::vector<Chunk>> parts(n);
        while (!todo.empty()) {
            const size_t m = todo.size();
            std::vector<LaneJob> jobs(m);
            uint64_t total = 0;
            for (size_t k = 0; k < m; k++) {
                jobs[k] = LaneJob{todo[k].s, total, todo[k].cap, (int32_t)(todo[k].id + 1)};
                total += todo[k].cap;
            }
            pools_.emplace_back(total * 2);
            uint16_t* pool = pools_.back().get<uint16_t>();
            d_jobs_.reserve(m * sizeof(LaneJob));
            d_work_.reserve(m * sizeof(seqd::Inflater));
            d_res_.reserve(m * sizeof(seqd::Outcome));
            HIP_OK(hipMemcpyAsync(d_jobs_.get(), jobs.data(), m * sizeof(LaneJob), hipMemcpyHostToDevice, dev_.st));
            HIP_OK(hipMemsetAsync(d_work_.get(), 0, m * sizeof(seqd::Inflater), dev_.st));
            HIP_OK(hipMemsetAsync(d_stored_count_.get(), 0, sizeof(unsigned), dev_.st));
            HIP_OK(hipMemsetAsync(d_copied_.get(), 0, 8, dev_.st));
            const gz::StoredList sl{d_stored_.get<gz::StoredCopy>(), d_stored_count_.get<unsigned>(), stored_cap_};
            clock_.time("k_rescue_decode", [&] {
                hipLaunchKernelGGL(k_rescue_decode, dim3((unsigned)((m + 63) / 64)), dim3(64), 0, dev_.st,
                                   (const uint8_t*)d_in_.get(), size_, (const Start*)d_targets_.get(), (int)n,
                                   (const LaneJob*)d_jobs_.get(), (int)m, pool, d_work_.get<seqd::Inflater>(),
                                   d_res_.get<seqd::Outcome>(), sl);
            });
            clock_.time("k_rescue_stored", [&] {
                hipLaunchKernelGGL(k_rescue_stored, dim3(grid_for(stored_cap_, 1, 4096)), dim3(256), 0, dev_.st,
                                   (const gz::StoredCopy*)sl.list, (const unsigned*)sl.count, sl.cap,
                                   (const uint8_t*)d_in_.get(), pool, d_copied_.get<unsigned long long>());
            });
            std::vector<seqd::Outcome> r(m);
            unsigned long long copied = 0;
            HIP_OK(hipMemcpyAsync(r.data(), d_res_.get(), m * sizeof(seqd::Outcome), hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipMemcpyAsync(&copied, d_copied_.get(), 8, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
            stored_bytes_ += copied;
            clock_.settle();
            std::vector<Call> next;
            for (size_t k = 0; k < m; k++) {
                const Call& c = todo[k];
                const seqd::Outcome& o = r[k];
                if (o.status == gz::ERR_OVERFLOW && c.cap < CALL_CAP) {
                    next.push_back(Call{c.id, c.s, (uint32_t)std::min<uint64_t>(4ull * c.cap, CALL_CAP)});
                    continue;
                }
                rsc::Decoded& d = res[c.id];
                const uint64_t base = d.len;
                d.blocks += o.blocks;
                d.symbols += o.symbols;
                if (o.last_block_bit != UINT64_MAX) {
                    d.last_block_bit = o.last_block_bit;
                    d.last_block_at = base + o.last_block_out;
                }
                d.len += o.out_len;
                parts[c.id].push_back(Chunk{pool + jobs[k].out, o.out_len});
                if (o.status == gz::ERR_OVERFLOW) {
                    next.push_back(Call{c.id, o.resume, c.cap});
                    continue;
                }
                d.status = o.status;
                d.end_bit = o.end_bit;
                d.landed = o.landed;
            }
            todo.swap(next);
        }
        for (size_t i = 0; i < n; i++) {
            res[i].id = (int)(first + i);
            outs_[first + i] = contiguous(parts[i], res[i].len);
        }
    }

    // on the host (rescue.cpp:661-750's loop), then to the GPU
    void decode_other(const Start& s0, bool stream_start, size_t t0, const std::vector<Start>& h, rsc::Decoded& d) {
        if (!host_scratch_) host_scratch_.reset(new uint16_t[CALL_CAP]);
        std::vector<uint16_t> out;
        Start s = s0;
        const Start* tg = h.data() + std::min(t0, h.size());
        const int nt = (int)(h.size() - std::min(t0, h.size()));
        d = rsc::Decoded{};
        for (;;) {
            const seqd::Outcome r =
                host_dec_.run(in_, size_, s, tg, nt, host_scratch_.get(), CALL_CAP, stream_start && out.empty());
            const uint64_t base = out.size();
            out.insert(out.end(), host_scratch_.get(), host_scratch_.get() + r.out_len);
            d.blocks += r.blocks;
            d.symbols += r.symbols;
            if (r.last_block_bit != UINT64_MAX) {
                d.last_block_bit = r.last_block_bit;
                d.last_block_at = base + r.last_block_out;
            }
            if (r.status == gz::ERR_OVERFLOW) {
                s = r.resume;
                stream_start = false;
                continue;
            }
            d.status = r.status;
            d.end_bit = r.end_bit;
            d.landed = r.landed;
            break;
        }
        d.len = out.size();
        d.id = (int)outs_.size();
        pools_.emplace_back(std::max<size_t>(2, out.size() * 2));
        HIP_OK(hipMemcpyAsync(pools_.back().get(), out.data(), out.size() * 2, hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        outs_.push_back(Chunk{pools_.back().get<uint16_t>(), d.len});
    }

    void resolve(const std::vector<rsc::Span>& spans, uint64_t total, rsc::Provenance& pv) {
        HIP_OK(hipSetDevice(dev_.device));
        auto store = std::make_shared<DevProv>();
        pv.total = total;
        pv.store = store;
        if (!total) return;
        std::vector<DevSpan> ds;
        std::vector<SpanChunk> chunks;
        for (const rsc::Span& sp : spans) {
            if (!sp.len) continue;
            // every link then points at an earlier byte of the output: d <= i, or d - i <= window <= out
            if (sp.id < 0 || (size_t)sp.id >= outs_.size() || outs_[sp.id].len < sp.len || sp.out + sp.len > total ||
                (sp.has_window && sp.window > sp.out))
                throw std::runtime_error("a stretch of the output outside its decode (internal error)");
            DevSpan d{};
            d.src = outs_[sp.id].p;
            d.out = sp.out;
            d.len = sp.len;
            d.window = sp.window;
            d.has_window = sp.has_window;
            d.var_base = sp.var_base;
            d.tail = sp.blank_tail ? (sp.len > gz::MAX_SYMBOL_OUT ? sp.len - gz::MAX_SYMBOL_OUT : 0) : sp.len;
            for (uint64_t a = 0; a < sp.len; a += PROV_CHUNK) chunks.push_back(SpanChunk{(uint32_t)ds.size(), 0, a});
            ds.push_back(d);
        }
        DevMem d_spans(std::max<size_t>(1, ds.size()) * sizeof(DevSpan));
        DevMem d_chunks(std::max<size_t>(1, chunks.size()) * sizeof(SpanChunk));
        DevMem ent(total * 8), left(8);
        HIP_OK(hipMemcpyAsync(d_spans.get(), ds.data(), ds.size() * sizeof(DevSpan), hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipMemcpyAsync(d_chunks.get(), chunks.data(), chunks.size() * sizeof(SpanChunk), hipMemcpyHostToDevice,
                              dev_.st));
        HIP_OK(hipMemsetAsync(ent.get(), 0, total * 8, dev_.st));
        if (!chunks.empty())
            clock_.time("k_prov_init", [&] {
                hipLaunchKernelGGL(k_prov_init, dim3((unsigned)chunks.size()), dim3(256), 0, dev_.st,
                                   (const DevSpan*)d_spans.get(), (const SpanChunk*)d_chunks.get(), ent.get<uint64_t>());
            });
        // each launch at least doubles every chain's reach: 64 launches would cover 2^64 bytes
        for (int round = 0;; round++) {
            if (round == 64) throw std::runtime_error("provenance links did not settle (internal error)");
            HIP_OK(hipMemsetAsync(left.get(), 0, 8, dev_.st));
            clock_.time("k_prov_jump", [&] {
                hipLaunchKernelGGL(k_prov_jump, dim3(grid_for(total, 256)), dim3(256), 0, dev_.st, ent.get<uint64_t>(),
                                   total, left.get<unsigned long long>());
            });
            unsigned long long still = 0;
            HIP_OK(hipMemcpyAsync(&still, left.get(), 8, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
            if (!still) break;
        }
        store->byte.reset(total);
        store->tag.reset(total * 4);
        clock_.time("k_prov_split", [&] {
            hipLaunchKernelGGL(k_prov_split, dim3(grid_for(total, 256)), dim3(256), 0, dev_.st, (const uint64_t*)ent.get(),
                               total, store->byte.get<uint8_t>(), store->tag.get<uint32_t>());
        });
        HIP_OK(hipStreamSynchronize(dev_.st));
        clock_.settle();
    }

    void summarize(const rsc::Provenance& pv, const std::vector<rsc::PieceRange>& pieces, std::vector<rsc::PieceFacts>& facts,
                   const std::vector<std::pair<uint64_t, uint64_t>>& crc_ranges, std::vector<uint32_t>& crcs) {
        HIP_OK(hipSetDevice(dev_.device));
        const DevProv& d = prov(pv);
        const size_t np = pieces.size();
        facts.assign(np, rsc::PieceFacts{});
        std::vector<Stretch> st(np);
        std::vector<Tile> tiles;
        for (size_t k = 0; k < np; k++) {
            check_stretch(pv, pieces[k].at, pieces[k].len);
            st[k] = Stretch{pieces[k].at, pieces[k].len};
            for (uint64_t a = 0; a < pieces[k].len; a += SUM_TILE) tiles.push_back(Tile{(uint32_t)k, 0, a});
        }
        if (!tiles.empty()) {
            DevMem d_st(np * sizeof(Stretch)), d_tiles(tiles.size() * sizeof(Tile)), d_bad(np * 4), d_v64(np * 8);
            HIP_OK(hipMemcpyAsync(d_st.get(), st.data(), np * sizeof(Stretch), hipMemcpyHostToDevice, dev_.st));
            HIP_OK(hipMemcpyAsync(d_tiles.get(), tiles.data(), tiles.size() * sizeof(Tile), hipMemcpyHostToDevice, dev_.st));
            HIP_OK(hipMemsetAsync(d_bad.get(), 0, np * 4, dev_.st));
            HIP_OK(hipMemsetAsync(d_v64.get(), 0xFF, np * 8, dev_.st));
            clock_.time("k_sum_pieces", [&] {
                hipLaunchKernelGGL(k_sum_pieces, dim3((unsigned)tiles.size()), dim3(256), 0, dev_.st,
                                   d.tag.get<const uint32_t>(), d_st.get<const Stretch>(), d_tiles.get<const Tile>(),
                                   d_bad.get<unsigned>(), d_v64.get<unsigned long long>());
            });
            std::vector<unsigned> bad(np);
            std::vector<unsigned long long> v64(np);
            HIP_OK(hipMemcpyAsync(bad.data(), d_bad.get(), np * 4, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipMemcpyAsync(v64.data(), d_v64.get(), np * 8, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
            for (size_t k = 0; k < np; k++) {
                facts[k].all_exact = !bad[k];
                facts[k].var64 = std::min<uint64_t>(v64[k], pieces[k].len);
            }
        }
        // the kept pieces' ranges with variables
        std::vector<size_t> kept;
        std::vector<Stretch> ks;
        for (size_t k = 0; k < np; k++)
            if (pieces[k].keep && pieces[k].len && !facts[k].all_exact) {
                kept.push_back(k);
                ks.push_back(st[k]);
            }
        if (!kept.empty()) var_runs(d, kept, ks, facts);
        crcs.resize(crc_ranges.size());
        for (size_t i = 0; i < crc_ranges.size(); i++) crcs[i] = crc_of(pv, crc_ranges[i].first, crc_ranges[i].second);
        clock_.settle();
    }

    // every byte of the stretches that copies a variable: its tag to the host, its place with the
    // cells around it kept on the device for places()
    void var_tags(const rsc::Provenance& pv, const std::vector<std::pair<uint64_t, uint64_t>>& stretches,
                  std::vector<uint32_t>& tags) {
        HIP_OK(hipSetDevice(dev_.device));
        tags.clear();
        nplaces_ = 0;
        const DevProv& d = prov(pv);
        std::vector<Stretch> st;
        std::vector<Tile> tiles;
        for (const auto& [out, len] : stretches) {
            check_stretch(pv, out, len);
            for (uint64_t a = 0; a < len; a += SCAN_TILE) tiles.push_back(Tile{(uint32_t)st.size(), 0, a});
            st.push_back(Stretch{out, len});
        }
        if (tiles.empty()) return;
        const size_t nt = tiles.size();
        DevMem d_st(st.size() * sizeof(Stretch)), d_tiles(nt * sizeof(Tile)), d_n(nt * 4), d_at(nt * 4);
        HIP_OK(hipMemcpyAsync(d_st.get(), st.data(), st.size() * sizeof(Stretch), hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipMemcpyAsync(d_tiles.get(), tiles.data(), nt * sizeof(Tile), hipMemcpyHostToDevice, dev_.st));
        clock_.time("k_var_places", [&] {
            hipLaunchKernelGGL(k_var_places<false>, dim3((unsigned)nt), dim3(64), 0, dev_.st, d.byte.get<const uint8_t>(),
                               d.tag.get<const uint32_t>(), d_st.get<const Stretch>(), d_tiles.get<const Tile>(),
                               d_n.get<unsigned>(), (const unsigned*)nullptr, (rsc::Place*)nullptr, (uint32_t*)nullptr);
        });
        std::vector<unsigned> n(nt), at(nt);
        HIP_OK(hipMemcpyAsync(n.data(), d_n.get(), nt * 4, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        uint64_t total = 0;
        for (size_t i = 0; i < nt; i++) {
            at[i] = (unsigned)total;
            total += n[i];
        }
        if (total >> 32) throw std::runtime_error("more than 4 Gi bytes copy lost data (internal limit)");
        if (total) {
            d_places_.reserve(total * sizeof(rsc::Place));
            DevMem d_tags(total * 4);
            HIP_OK(hipMemcpyAsync(d_at.get(), at.data(), nt * 4, hipMemcpyHostToDevice, dev_.st));
            clock_.time("k_var_places", [&] {
                hipLaunchKernelGGL(k_var_places<true>, dim3((unsigned)nt), dim3(64), 0, dev_.st, d.byte.get<const uint8_t>(),
                                   d.tag.get<const uint32_t>(), d_st.get<const Stretch>(), d_tiles.get<const Tile>(),
                                   (unsigned*)nullptr, d_at.get<const unsigned>(), d_places_.get<rsc::Place>(),
                                   d_tags.get<uint32_t>());
            });
            tags.resize(total);
            HIP_OK(hipMemcpyAsync(tags.data(), d_tags.get(), total * 4, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
        }
        nplaces_ = total;
        clock_.settle();
    }

    // the places of the last var_tags() call at the picked indices (ascending)
    void places(const std::vector<uint64_t>& picked, std::vector<rsc::Place>& out) {
        out.resize(picked.size());
        if (picked.empty()) return;
        std::vector<uint32_t> idx(picked.size());
        for (size_t j = 0; j < picked.size(); j++) {
            if (picked[j] >= nplaces_) throw std::runtime_error("a place that was not found (internal error)");
            idx[j] = (uint32_t)picked[j];
        }
        DevMem d_idx(idx.size() * 4), d_out(idx.size() * sizeof(rsc::Place));
        HIP_OK(hipMemcpyAsync(d_idx.get(), idx.data(), idx.size() * 4, hipMemcpyHostToDevice, dev_.st));
        clock_.time("k_pick_places", [&] {
            hipLaunchKernelGGL(k_pick_places, dim3(grid_for(idx.size(), 256, 4096)), dim3(256), 0, dev_.st,
                               d_places_.get<const rsc::Place>(), d_idx.get<const uint32_t>(), (unsigned)idx.size(),
                               d_out.get<rsc::Place>());
        });
        HIP_OK(hipMemcpyAsync(out.data(), d_out.get(), idx.size() * sizeof(rsc::Place), hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        clock_.settle();
    }

    void fetch(const rsc::Provenance& pv, uint64_t at, uint64_t len, uint8_t* byte, uint32_t* tag) {
        check_stretch(pv, at, len);
        if (!len) return;
        const DevProv& d = prov(pv);
        HIP_OK(hipMemcpyAsync(byte, d.byte.get<uint8_t>() + at, len, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipMemcpyAsync(tag, d.tag.get<uint32_t>() + at, len * 4, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
    }

    // The output, a window of device memory at a time, then to the host through two pinned
    // buffers in turn, to the sink a megabyte at a time (as the host emitter cuts it); after a sink
    // that stops, the windows are still made for their counts, and no more is copied
    int emit(const rsc::Provenance& pv, const std::vector<rsc::EmitPiece>& pieces, const rsc::Inferred& value, uint8_t fill,
             gr_sink sink, void* user, std::vector<rsc::EmitCounts>& counts, uint64_t& emitted) {
        HIP_OK(hipSetDevice(dev_.device));
        const size_t np = pieces.size();
        counts.assign(np, rsc::EmitCounts{});
        std::vector<EPiece> ep(np);
        uint64_t total = 0;
        for (size_t k = 0; k < np; k++) {
            const rsc::EmitPiece& p = pieces[k];
            check_stretch(pv, p.at, std::min(p.decoded, p.length));
            ep[k] = EPiece{p.at, p.decoded, p.length, total, p.suspect ? 1u : 0u, 0};
            total += p.length;
        }
        emitted = total;
        if (!total) return GR_OK;
        const DevProv& d = prov(pv);
        // the inferred values, sorted by variable
        std::vector<std::pair<uint32_t, uint8_t>> iv;
        if (value.vars)
            for (size_t i = 0; i < value.vars->list.size(); i++)
                if (value.value[i] >= 0) iv.push_back({value.vars->list[i].var, (uint8_t)value.value[i]});
        std::sort(iv.begin(), iv.end());
        std::vector<uint32_t> vars(iv.size());
        std::vector<uint8_t> vals(iv.size());
        for (size_t i = 0; i < iv.size(); i++) vars[i] = iv[i].first, vals[i] = iv[i].second;
        DevMem d_p(np * sizeof(EPiece)), d_c(np * sizeof(ECounts)), d_vars(std::max<size_t>(4, vars.size() * 4)),
            d_vals(std::max<size_t>(1, vals.size()));
        HIP_OK(hipMemcpyAsync(d_p.get(), ep.data(), np * sizeof(EPiece), hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipMemsetAsync(d_c.get(), 0, np * sizeof(ECounts), dev_.st));
        if (!vars.empty()) {
            HIP_OK(hipMemcpyAsync(d_vars.get(), vars.data(), vars.size() * 4, hipMemcpyHostToDevice, dev_.st));
            HIP_OK(hipMemcpyAsync(d_vals.get(), vals.data(), vals.size(), hipMemcpyHostToDevice, dev_.st));
        }
        constexpr uint64_t WINDOW = 512ull << 20, SLICE = 8ull << 20, CHUNK = rsc::Emitter::CHUNK;
        static_assert(WINDOW % SLICE == 0 && SLICE % CHUNK == 0 && WINDOW % EMIT_TILE == 0, "window layout");
        const uint64_t wbytes = std::min(WINDOW, (total + EMIT_TILE - 1) / EMIT_TILE * EMIT_TILE);
        DevMem out(wbytes), mask(wbytes);
        Staging& stg = dev_.staging(SLICE);
        int rc = GR_OK;
        for (uint64_t w0 = 0; w0 < total; w0 += WINDOW) {
            const uint64_t wn = std::min(WINDOW, total - w0);
            const uint64_t t0 = w0 / EMIT_TILE, nt = (wn + EMIT_TILE - 1) / EMIT_TILE;
            clock_.time("k_emit", [&] {
                hipLaunchKernelGGL(k_emit, dim3((unsigned)nt), dim3(256), 0, dev_.st, d.byte.get<const uint8_t>(),
                                   d.tag.get<const uint32_t>(), d_p.get<const EPiece>(), (int)np, total, t0, w0,
                                   d_vars.get<const uint32_t>(), d_vals.get<const uint8_t>(), (int)vars.size(), fill,
                                   out.get<uint8_t>(), mask.get<uint8_t>(), d_c.get<ECounts>());
            });
            if (rc != GR_OK || !sink) continue;
            // slices of the window through the two staging buffers: the copy of the next runs while
            // the sink takes this one
            const uint64_t ns = (wn + SLICE - 1) / SLICE;
            auto issue = [&](uint64_t j) {
                const uint64_t a = j * SLICE, n = std::min(SLICE, wn - a);
                Staging::Buf& b = stg.buf[j & 1];
                HIP_OK(hipMemcpyAsync(b.out, out.get<uint8_t>() + a, n, hipMemcpyDeviceToHost, dev_.st));
                HIP_OK(hipMemcpyAsync(b.mask, mask.get<uint8_t>() + a, n, hipMemcpyDeviceToHost, dev_.st));
                HIP_OK(hipEventRecord(b.done, dev_.st));
            };
            issue(0);
            if (ns > 1) issue(1);
            for (uint64_t j = 0; j < ns; j++) {
                Staging::Buf& b = stg.buf[j & 1];
                HIP_OK(hipEventSynchronize(b.done));
                const uint64_t n = std::min(SLICE, wn - j * SLICE);
                for (uint64_t c = 0; c < n && rc == GR_OK; c += CHUNK)
                    if (sink(user, b.out + c, b.mask + c, (size_t)std::min(CHUNK, n - c))) rc = GR_ERR_ABORTED;
                if (j + 2 < ns) {
                    if (rc != GR_OK) break;
                    issue(j + 2);
                }
            }
            HIP_OK(hipStreamSynchronize(dev_.st));
        }
        std::vector<ECounts> c(np);
        HIP_OK(hipMemcpyAsync(c.data(), d_c.get(), np * sizeof(ECounts), hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        clock_.settle();
        for (size_t k = 0; k < np; k++) counts[k] = rsc::EmitCounts{c[k].known, c[k].exact, c[k].suspect, c[k].unknown};
        return rc;
    }

  private:
    struct Chunk {
        const uint16_t* p = nullptr;
        uint64_t len = 0;
    };
    // a pass's provenance in device memory: a byte and a 32-bit tag an output byte
    struct DevProv {
        DevMem byte, tag;
    };
    static const DevProv& prov(const rsc::Provenance& pv) { return *static_cast<const DevProv*>(pv.store.get()); }
    static void check_stretch(const rsc::Provenance& pv, uint64_t at, uint64_t len) {
        if (at > pv.total || len > pv.total - at) throw std::runtime_error("a stretch outside the pass's output (internal error)");
    }

    // the ranges of variables in the kept pieces (facts[kept[i]].runs), in order
    void var_runs(const DevProv& d, const std::vector<size_t>& kept, const std::vector<Stretch>& ks,
                  std::vector<rsc::PieceFacts>& facts) {
        std::vector<Tile> tiles;
        std::vector<size_t> first_tile(kept.size() + 1);
        for (size_t i = 0; i < kept.size(); i++) {
            first_tile[i] = tiles.size();
            for (uint64_t a = 0; a < ks[i].len; a += SCAN_TILE) tiles.push_back(Tile{(uint32_t)i, 0, a});
        }
        first_tile[kept.size()] = tiles.size();
        const size_t nt = tiles.size();
        DevMem d_st(ks.size() * sizeof(Stretch)), d_tiles(nt * sizeof(Tile)), d_ns(nt * 4), d_ne(nt * 4), d_as(nt * 4),
            d_ae(nt * 4);
        HIP_OK(hipMemcpyAsync(d_st.get(), ks.data(), ks.size() * sizeof(Stretch), hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipMemcpyAsync(d_tiles.get(), tiles.data(), nt * sizeof(Tile), hipMemcpyHostToDevice, dev_.st));
        clock_.time("k_var_edges", [&] {
            hipLaunchKernelGGL(k_var_edges<false>, dim3((unsigned)nt), dim3(64), 0, dev_.st, d.tag.get<const uint32_t>(),
                               d_st.get<const Stretch>(), d_tiles.get<const Tile>(), d_ns.get<unsigned>(), d_ne.get<unsigned>(),
                               (const unsigned*)nullptr, (const unsigned*)nullptr, (uint64_t*)nullptr, (uint64_t*)nullptr);
        });
        std::vector<unsigned> ns(nt), ne(nt), as(nt), ae(nt);
        HIP_OK(hipMemcpyAsync(ns.data(), d_ns.get(), nt * 4, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipMemcpyAsync(ne.data(), d_ne.get(), nt * 4, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        uint64_t ts = 0, te = 0;
        for (size_t i = 0; i < nt; i++) {
            as[i] = (unsigned)ts;
            ae[i] = (unsigned)te;
            ts += ns[i];
            te += ne[i];
        }
        if (ts != te || (ts >> 32)) throw std::runtime_error("ranges of variables that do not pair up (internal error)");
        if (!ts) return;
        DevMem d_s(ts * 8), d_e(ts * 8);
        HIP_OK(hipMemcpyAsync(d_as.get(), as.data(), nt * 4, hipMemcpyHostToDevice, dev_.st));
        HIP_OK(hipMemcpyAsync(d_ae.get(), ae.data(), nt * 4, hipMemcpyHostToDevice, dev_.st));
        clock_.time("k_var_edges", [&] {
            hipLaunchKernelGGL(k_var_edges<true>, dim3((unsigned)nt), dim3(64), 0, dev_.st, d.tag.get<const uint32_t>(),
                               d_st.get<const Stretch>(), d_tiles.get<const Tile>(), (unsigned*)nullptr, (unsigned*)nullptr,
                               d_as.get<const unsigned>(), d_ae.get<const unsigned>(), d_s.get<uint64_t>(), d_e.get<uint64_t>());
        });
        std::vector<uint64_t> starts(ts), ends(ts);
        HIP_OK(hipMemcpyAsync(starts.data(), d_s.get(), ts * 8, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipMemcpyAsync(ends.data(), d_e.get(), ts * 8, hipMemcpyDeviceToHost, dev_.st));
        HIP_OK(hipStreamSynchronize(dev_.st));
        for (size_t i = 0; i < kept.size(); i++) {
            const uint64_t a = first_tile[i] < nt ? as[first_tile[i]] : ts;
            const uint64_t b = first_tile[i + 1] < nt ? as[first_tile[i + 1]] : ts;
            auto& runs = facts[kept[i]].runs;
            for (uint64_t r = a; r < b; r++) runs.push_back({starts[r], ends[r]});
        }
    }

    // gzip's CRC-32 of the provenance's bytes [a, e): 64 KiB chunks on the GPU, combined on the
    // host with the rest
    uint32_t crc_of(const rsc::Provenance& pv, uint64_t a, uint64_t e) {
        check_stretch(pv, a, e - a);
        const DevProv& d = prov(pv);
        const uint64_t n = (e - a) / CRC_CHUNK, tail = (e - a) % CRC_CHUNK;
        uint32_t crc = 0;
        if (n) {
            CrcShifts sh;
            for (int j = 0; j < 64; j++) sh.k[j] = hostcrc::x8nmodp(1024ull * j);
            DevMem d_out(n * 4);
            clock_.time("k_crc_chunks", [&] {
                hipLaunchKernelGGL(k_crc_chunks, dim3(grid_for(n, 1, 65536)), dim3(64), 0, dev_.st, d.byte.get<const uint8_t>(),
                                   a, n, sh, d_out.get<uint32_t>());
            });
            std::vector<uint32_t> c(n);
            HIP_OK(hipMemcpyAsync(c.data(), d_out.get(), n * 4, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
            const uint32_t shift = hostcrc::x8nmodp(CRC_CHUNK);
            crc = c[0];
            for (uint64_t i = 1; i < n; i++) crc = hostcrc::multmodp(shift, crc) ^ c[i];
        }
        if (tail) {
            std::vector<uint8_t> t(tail);
            HIP_OK(hipMemcpyAsync(t.data(), d.byte.get<uint8_t>() + a + n * CRC_CHUNK, tail, hipMemcpyDeviceToHost, dev_.st));
            HIP_OK(hipStreamSynchronize(dev_.st));
            const uint32_t ct = hostcrc::crc32(0, t.data(), tail);
            crc = n ? hostcrc::combine(crc, ct, tail) : ct;
        }
        return crc;
    }

    // the header search over in[0, n), headers from bit `from` (>= 1) on
    void search(const uint8_t* in, uint64_t n, uint64_t from, std::vector<uint64_t>& out) {
        uint64_t cap = n / 64 + 4096, count = 0;
        for (;;) {
            out.resize(cap);
            const int rc = gi_rescue_headers(dev_.finder, in, n, std::max<uint64_t>(from, 1), out.data(), cap, &count);
            if (rc != GI_OK) throw GpuFault(std::string("gi_rescue_headers: ") + gi_last_error(dev_.finder));
            if (count <= cap) break;
            cap = count;
        }
        out.resize(count);
    }

    // a decode's calls as one stretch of device memory
    Chunk contiguous(const std::vector<Chunk>& parts, uint64_t len) {
        if (parts.size() == 1) return parts[0];
        pools_.emplace_back(std::max<uint64_t>(2, len * 2));
        uint16_t* p = pools_.back().get<uint16_t>();
        uint64_t at = 0;
        for (const Chunk& c : parts) {
            if (c.len)
                HIP_OK(hipMemcpyAsync(p + at, c.p, c.len * 2, hipMemcpyDeviceToDevice, dev_.st));
            at += c.len;
        }
        HIP_OK(hipStreamSynchronize(dev_.st));
        return Chunk{p, len};
    }

    Device& dev_;
    const uint8_t* in_ = nullptr;
    uint64_t size_ = 0;
    DevMem d_in_, d_targets_, d_jobs_, d_work_, d_res_, d_stored_, d_stored_count_, d_copied_;
    DevMem d_places_;  // the places of the last var_tags() call
    uint64_t nplaces_ = 0;
    unsigned stored_cap_ = 0;
    uint64_t stored_bytes_ = 0;
    KernelClock clock_;
    std::vector<DevMem> pools_;  // every decode's symbols
    std::vector<Chunk> outs_;    // by id
    seqd::Inflater host_dec_;
    std::unique_ptr<uint16_t[]> host_scratch_;
};

thread_local std::string last_error;

}  // namespace

struct gr_result : rsc::RescueResult {
    std::vector<gzr_kernel_time> kernels;  // GPU time per kernel
    uint64_t stored_bytes = 0;
};

namespace {
int rescue_entry(const uint8_t* in, uint64_t size, bool gzip, uint64_t expected_size, bool has_crc, uint32_t expected_crc,
                 const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    if (result) *result = nullptr;
    last_error.clear();
    if (!in && size) return GR_ERR_ARG;
    Device& dev = the_device();
    std::lock_guard<std::mutex> lock(dev.mu);
    try {
        dev.open();
        GpuRescue be(dev);
        std::unique_ptr<gr_result> res(new gr_result);
        const int rc =
            rsc::run_rescue(be, in, size, gzip, expected_size, has_crc, expected_crc, rsc::read_options(opt), sink, user, *res);
        res->kernels = be.kernel_times();
        res->stored_bytes = be.stored_bytes();
        if (rc != GR_OK && rc != GR_ERR_ABORTED) return rc;
        if (result) *result = res.release();
        return rc;
    } catch (const std::bad_alloc&) {
        last_error = "out of host memory";
        return GR_ERR_MEMORY;
    } catch (const GpuFault& e) {
        last_error = std::string("GPU: ") + e.what();
        return GR_ERR_INTERNAL;
    } catch (const std::exception& e) {
        last_error = e.what();
        return GR_ERR_INTERNAL;
    }
}
}  // namespace

extern "C" {

GR_API uint32_t gr_abi_version(void) { return GR_ABI_VERSION; }

GR_API int gr_rescue(const uint8_t* in, uint64_t size, const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    return rescue_entry(in, size, true, UINT64_MAX, false, 0, opt, sink, user, result);
}

GR_API int gr_rescue_raw(const uint8_t* in, uint64_t size, uint64_t expected_size, int has_crc, uint32_t expected_crc,
                         const gr_options* opt, gr_sink sink, void* user, gr_result** result) {
    return rescue_entry(in, size, false, expected_size, has_crc != 0, expected_crc, opt, sink, user, result);
}

GR_API const gr_stats* gr_result_stats(const gr_result* r) { return r ? &r->stats : nullptr; }

GR_API uint32_t gr_result_pieces(const gr_result* r, const gr_piece** pieces) {
    if (!r) return 0;
    if (pieces) *pieces = r->pieces.data();
    return (uint32_t)r->pieces.size();
}

GR_API uint32_t gr_result_notes(const gr_result* r, const char* const** notes) {
    if (!r) return 0;
    if (notes) *notes = r->note_ptrs.data();
    return (uint32_t)r->note_ptrs.size();
}

GR_API void gr_free(gr_result* r) { delete r; }

GR_API uint32_t gzr_result_phases(const gr_result* r, const char** names, double* seconds, uint32_t cap) {
    if (!r) return 0;
    for (uint32_t i = 0; i < cap && i < r->phases.size(); i++) {
        if (names) names[i] = r->phases[i].first.c_str();
        if (seconds) seconds[i] = r->phases[i].second;
    }
    return (uint32_t)r->phases.size();
}

GR_API int gzr_result_kernels(const gr_result* r, gzr_kernel_time* out, int max, uint64_t* stored_bytes) {
    if (!r) return 0;
    for (int i = 0; i < max && i < (int)r->kernels.size(); i++) out[i] = r->kernels[i];
    if (stored_bytes) *stored_bytes = r->stored_bytes;
    return (int)r->kernels.size();
}

GR_API const char* gzr_last_error(void) { return last_error.c_str(); }

GR_API int gzr_set_device(int device) {
    if (device < 0) return GR_ERR_ARG;
    Device& dev = the_device();
    std::lock_guard<std::mutex> lock(dev.mu);
    if (device != dev.device) {
        dev.close();
        dev.device = device;
    }
    return GR_OK;
}

}  // extern "C"
