// IEEE binary16 conversions of the rung trace (src/creature_kernel.rs:221-274).
#include "host_common.hpp"

namespace evo {

uint32_t f32_to_f16(float v) {
    const uint32_t bits = as_uint(v);
    const uint32_t sign = (bits >> 16) & 0x8000u;
    const int32_t exponent = static_cast<int32_t>((bits >> 23) & 0xffu);
    const uint32_t mantissa = bits & 0x7fffffu;
    if (exponent == 0xff) { return sign | 0x7c00u | (mantissa != 0u ? 0x200u : 0u); }
    const int32_t e = exponent - 127 + 15;
    if (e >= 31) { return sign | 0x7c00u; }
    if (e <= 0) {
        if (e < -10) { return sign; }
        const uint32_t m = mantissa | 0x800000u;
        const uint32_t shift = static_cast<uint32_t>(14 - e);
        const uint32_t half = 1u << (shift - 1u);
        const uint32_t rest = m & ((1u << shift) - 1u);
        uint32_t out = m >> shift;
        if (rest > half || (rest == half && (out & 1u) == 1u)) { out += 1u; }
        return sign | out;
    }
    const uint32_t rest = mantissa & 0x1fffu;
    uint32_t out = (static_cast<uint32_t>(e) << 10) | (mantissa >> 13);
    if (rest > 0x1000u || (rest == 0x1000u && (out & 1u) == 1u)) { out += 1u; }
    return sign | out;
}

float f16_to_f32(uint32_t h) {
    const uint32_t sign = ((h >> 15) & 1u) << 31;
    const uint32_t exponent = (h >> 10) & 0x1fu;
    const uint32_t mantissa = h & 0x3ffu;
    uint32_t bits;
    if (exponent == 0u && mantissa == 0u) {
        bits = sign;
    } else if (exponent == 0u) {
        uint32_t shift = 0u, m = mantissa;
        while ((m & 0x400u) == 0u) { m <<= 1; shift++; }
        bits = sign | ((113u - shift) << 23) | ((m & 0x3ffu) << 13);
    } else if (exponent == 31u && mantissa == 0u) {
        bits = sign | 0x7f800000u;
    } else if (exponent == 31u) {
        bits = sign | 0x7fc00000u | (mantissa << 13);
    } else {
        bits = sign | ((exponent + 112u) << 23) | (mantissa << 13);
    }
    return as_float(bits);
}

}  // namespace evo
