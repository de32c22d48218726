// ref_shim.cpp - a C entry layer over the reference's own host-compiled decoder
// (gpuinflate/inflate.cuh, gi::Decoder), built from the reference sources where they lie into
// oracle/_ref/libgiref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: it pins the oracle restatement (inflate_oracle.c) and serves as the
// "reference" CPU baseline in bench.py. Nothing in the product loads it.
//
// ref_inflate follows gi_test.cpp:47-91 (`reference`): decode from bit 0 with no targets,
// continue after a full buffer, then resolve the back-references in order.
#include "inflate.cuh"

#include <stdint.h>
#include <string.h>

#include <vector>

extern "C" {

struct ref_start {
    uint64_t bit, header;
    uint8_t mode;
    int8_t final;
};

struct ref_result {
    uint64_t end_bit, eob_bit;
    ref_start resume;
    uint32_t eob_out, out_len;
    int32_t status, landed, end_final;
    uint32_t blocks;
    uint64_t last_block_bit;
    uint32_t last_block_out, symbols;
};

__attribute__((visibility("default"))) int ref_plausible_dynamic_header(const uint8_t* in, uint64_t size, uint64_t bit) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    return gi::plausible_dynamic_header(pad.data(), size, bit) ? 1 : 0;
}

__attribute__((visibility("default"))) int ref_header_prefilter128(uint64_t v1, uint64_t v2) {
    return gi::header_prefilter128(v1, v2) ? 1 : 0;
}

__attribute__((visibility("default"))) int ref_header_prefilter(uint64_t v) { return gi::header_prefilter(v) ? 1 : 0; }

// Decoder::run on a padded copy of the input
__attribute__((visibility("default"))) void ref_run(const uint8_t* in, uint64_t size, const ref_start* s,
                                                     const ref_start* targets, int ntargets, uint16_t* out, uint32_t cap,
                                                     int stream_start, uint64_t stop, ref_result* res) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    static thread_local gi::Decoder* dec = nullptr;
    static thread_local uint16_t luts[gi::LUT_U16];
    if (!dec) dec = new gi::Decoder;
    dec->attach(luts);
    std::vector<gi::Start> tg(ntargets);
    for (int i = 0; i < ntargets; i++) tg[i] = {targets[i].bit, targets[i].header, targets[i].mode, targets[i].final};
    gi::Start st{s->bit, s->header, s->mode, s->final};
    gi::Result r = dec->run(pad.data(), size, st, ntargets ? tg.data() : nullptr, ntargets, out, cap, stream_start != 0, stop);
    res->end_bit = r.end_bit;
    res->eob_bit = r.eob_bit;
    res->resume = {r.resume.bit, r.resume.header, r.resume.mode, r.resume.final};
    res->eob_out = r.eob_out;
    res->out_len = r.out_len;
    res->status = r.status;
    res->landed = r.landed;
    res->end_final = r.end_final;
    res->blocks = r.blocks;
    res->last_block_bit = r.last_block_bit;
    res->last_block_out = r.last_block_out;
    res->symbols = r.symbols;
}

// gi_test.cpp:47-91. Returns 0 ok, 1 data, 2 truncated; *too_big when more than max_out.
__attribute__((visibility("default"))) int ref_inflate(const uint8_t* in, uint64_t size, uint8_t* out, uint64_t max_out,
                                                        uint64_t* out_len, uint64_t* in_used, int* too_big) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    std::vector<uint16_t> o16(max_out + 2 * gi::MAX_SYMBOL_OUT);
    static thread_local gi::Decoder* dec = nullptr;
    static thread_local uint16_t luts[gi::LUT_U16];
    if (!dec) dec = new gi::Decoder;
    dec->attach(luts);
    *out_len = 0;
    *in_used = 0;
    *too_big = 0;
    gi::Start s{0, 0, gi::AT_BLOCK, -1};
    uint64_t op = 0;
    for (;;) {
        uint64_t room = o16.size() - op;
        gi::Result r = dec->run(pad.data(), size, s, nullptr, 0, o16.data() + op,
                                (uint32_t)(room > 0xFFFFFFFFull ? 0xFFFFFFFFull : room), op == 0);
        op += r.out_len;
        if (r.status == gi::ERR_OVERFLOW) {
            if (o16.size() - op < 64 * 1024) {
                *too_big = 1;
                return 0;
            }
            s = r.resume;
            continue;
        }
        if (r.status != gi::OK) return r.status == gi::ERR_INPUT ? 2 : 1;
        *in_used = (r.end_bit + 7) / 8;
        break;
    }
    if (op > max_out) {
        *too_big = 1;
        return 0;
    }
    for (uint64_t i = 0; i < op; i++) {
        uint16_t v = o16[i];
        if (v <= gi::REF) {
            out[i] = (uint8_t)v;
        } else if (v - gi::REF > i) {
            *in_used = 0;
            return 1;
        } else {
            out[i] = out[i - (v - gi::REF)];
        }
    }
    *out_len = op;
    return 0;
}

}  // extern "C"

extern "C" {
// Decoder::probe (inflate.cuh:591-648) at one point
__attribute__((visibility("default"))) void ref_probe(const uint8_t* in, uint64_t size, uint8_t mode, uint64_t header,
                                                       uint64_t from_bit, uint64_t at_bit, ref_start* out) {
    std::vector<uint8_t> pad((size + 3) / 4 * 4 + 64, 0);
    memcpy(pad.data(), in, size);
    static thread_local gi::Decoder* dec = nullptr;
    static thread_local uint16_t luts[gi::LUT_U16];
    if (!dec) dec = new gi::Decoder;
    dec->attach(luts);
    gi::Start s = dec->probe(pad.data(), size, mode, header, from_bit, at_bit);
    *out = {s.bit, s.header, s.mode, s.final};
}
}
