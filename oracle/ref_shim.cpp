// ref_shim.cpp -- C entry points over the reference's own header-only host
// code, compiled from /root/reference/src by oracle/Makefile into
// oracle/_ref/libamref.so. TEST INFRASTRUCTURE ONLY: used by tests/ to pin
// the oracle's restatement against the reference itself.
//
// Only the two reference headers that build with this image's toolchain are
// used: src/fast_prng.cuh and src/bit_iterator.cuh (hipcc supplies
// __host__/__device__). The rest of the reference includes the CUDA runtime
// (src/cuda_try.cuh, src/cuda_time.cuh) and is not buildable here.
#include <cstddef>
#include <cstdint>
#include "fast_prng.cuh"
#include "bit_iterator.cuh"

extern "C" {

// draws skip..skip+n-1 of fast_prng(seed) (src/fast_prng.cuh:21-30)
void amref_prng_fill(uint64_t seed, uint64_t skip, uint32_t* out, size_t n)
{
    fast_prng rng(seed);
    for (uint64_t i = 0; i < skip; i++) rng.rand();
    for (size_t i = 0; i < n; i++) out[i] = rng.rand();
}

// out[i] = bit i of the byte mask read through bit_iterator: alternately via
// operator[] and via post-increment + operator* (src/bit_iterator.cuh:13-21,45-49)
void amref_bit_iterator_read(uint8_t* bytes, size_t nbits, uint8_t* out_index, uint8_t* out_walk)
{
    bit_iterator it{bytes};
    for (size_t i = 0; i < nbits; i++) out_index[i] = it[i] ? 1 : 0;
    bit_iterator walk{bytes, 0};
    for (size_t i = 0; i < nbits; i++) {
        out_walk[i] = *walk ? 1 : 0;
        walk++;
    }
}
}
