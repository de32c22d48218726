// rescue_shim.cpp - a C entry layer over the reference's own rescue library
// (gpuinflate/rescue.cpp, which includes inflate.cuh), built from the reference sources where
// they lie into oracle/_ref/libgrref.so by oracle/Makefile.
//
// TEST INFRASTRUCTURE ONLY: it pins the GPU rescue steps against the reference. Nothing in the
// product loads it. Including rescue.cpp in this one translation unit also exports its gr_* C
// interface (GR_BUILD), so the whole reference rescue can be run from the tests.
#define GR_BUILD
#include "rescue.cpp"

extern "C" {

// Rescuer::find_headers (rescue.cpp:597-626): every confirmed dynamic-block header in
// [from, nbits - 128), in bit order; the first `cap` go to out. Returns how many there are.
__attribute__((visibility("default"))) uint64_t grref_find_headers(const uint8_t* in, uint64_t size, uint64_t from,
                                                                   uint64_t* out, uint64_t cap) {
    std::vector<uint8_t> buf(size + 64, 0);  // rescue_common's zero padding: the readers look ahead
    if (size) memcpy(buf.data(), in, size);
    Rescuer rs(buf.data(), size);
    std::vector<gi::Start> h = rs.find_headers(from);
    for (uint64_t i = 0; i < h.size() && i < cap; i++) out[i] = h[i].bit;
    return h.size();
}

// verify_at (rescue.cpp:549-554) at one bit: 1 when decoding from a block start there goes on
// cleanly for long enough to trust it
__attribute__((visibility("default"))) int grref_verify_at(const uint8_t* in, uint64_t size, uint64_t bit) {
    std::vector<uint8_t> buf(size + 64, 0);
    if (size) memcpy(buf.data(), in, size);
    Verifier vf;
    gi::Start s = {bit, bit, gi::AT_BLOCK, -1};
    return verify_at(buf.data(), size, s, vf.dec, vf.scratch.data()) ? 1 : 0;
}

}  // extern "C"
