// amethyst.hpp -- header-only C++ facade over the C ABI (amethyst.h) that keeps
// the reference's templated operator surface: compile-time <bits, outbits,
// op> wrappers named after the reference's host wrappers (src/lib.cuh:4-46)
// and kernels (src/kernels/*), plus bit_iterator / slab_iterator views
// (src/bit_iterator.cuh, src/slab_iterator.cuh). Every call forwards to the
// gfx950 library; nothing here computes on the host except the iterators.
//
// Differences from the reference, all deliberate:
//   * errors come back as am_status codes (the reference aborts via CUDA_TRY);
//   * every wrapper takes a stream (default: the null stream);
//   * block/grid sizes are not template parameters -- launch shapes are the
//     library's business (persistent, occupancy-sized grids);
//   * compressstore needs a workspace (see Workspace) and reports its count
//     on the device; compressstore_slabs_count() synchronises and returns it.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <type_traits>

#include "amethyst.h"

namespace amethyst {

// ---------------------------------------------------------------- widths

// tight_type<bits> (src/common_code.cuh:11-19): smallest unsigned container.
template <int bits>
struct tight_type {
    static constexpr size_t bytecount = (bits + 7) / 8;
    using type = typename std::conditional<
        (bytecount > 4), uint64_t,
        typename std::conditional<(bytecount > 2), uint32_t,
                                  typename std::conditional<(bytecount > 1), uint16_t, uint8_t>::type>::type>::type;
};

constexpr uint64_t gen_bitmask(int bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull); }
constexpr int add_outbits(int bits) { return bits >= 64 ? 64 : bits + 1; }
constexpr int mul_outbits(int bits) { return bits * 2 >= 64 ? 64 : bits * 2; }
constexpr int mod_outbits(int bits) { return bits; }

// ---------------------------------------------------------------- iterators

// LSB-first bool view of a byte mask: element i is bit i%8 of byte i/8.
struct bit_iterator {
    uint8_t* byte_ptr;
    size_t idx;
    bit_iterator(uint8_t* p, size_t i = 0) : byte_ptr(p), idx(i) {}
    bool operator*() const { return (byte_ptr[idx >> 3] >> (idx & 7)) & 1u; }
    bool operator[](size_t i) const { return bit_iterator(byte_ptr, idx + i).operator*(); }
    bit_iterator operator+(size_t i) const { return bit_iterator(byte_ptr, idx + i); }
    bit_iterator operator-(size_t i) const { return bit_iterator(byte_ptr, idx - i); }
    bit_iterator& operator++() { ++idx; return *this; }
    bit_iterator& operator--() { --idx; return *this; }
    bit_iterator operator++(int) { bit_iterator c = *this; ++idx; return c; }
    bit_iterator operator--(int) { bit_iterator c = *this; --idx; return c; }
};

// Masked proxy for element `index` of a tight array (the reference's
// sentinel<bits>): reads return the value masked to `bits`, writes store it
// masked.
template <int bits>
struct sentinel {
    using raw_t = typename tight_type<bits>::type;
    static constexpr raw_t bitmask = (raw_t)gen_bitmask(bits);
    raw_t* data;
    size_t index;
    sentinel(void* d, size_t i = 0) : data((raw_t*)d), index(i) {}
    operator raw_t() const { return data[index] & bitmask; }
    sentinel& operator=(raw_t v)
    {
        data[index] = v & bitmask;
        return *this;
    }
    sentinel& operator=(const sentinel& o) { return *this = (raw_t)o; }
    bool operator==(const sentinel& o) const { return (raw_t)*this == (raw_t)o; }
    bool operator<(const sentinel& o) const { return (raw_t)*this < (raw_t)o; }
};

template <int bits>
struct slab_iterator {
    using iterator_category = std::random_access_iterator_tag;
    using value_type = typename tight_type<bits>::type;
    using difference_type = std::ptrdiff_t;
    using reference = sentinel<bits>;
    void* data;
    size_t index;
    slab_iterator(void* d, size_t i = 0) : data(d), index(i) {}
    sentinel<bits> operator*() const { return sentinel<bits>(data, index); }
    sentinel<bits> operator[](size_t i) const { return sentinel<bits>(data, index + i); }
    slab_iterator operator+(size_t i) const { return slab_iterator(data, index + i); }
    slab_iterator operator-(size_t i) const { return slab_iterator(data, index - i); }
    difference_type operator-(const slab_iterator& o) const { return (difference_type)index - (difference_type)o.index; }
    slab_iterator& operator++() { ++index; return *this; }
    slab_iterator& operator--() { --index; return *this; }
};

// ---------------------------------------------------------------- columns

template <int bits>
inline am_column slabs(const uint64_t* d, size_t n) { return am_column{(void*)d, n, bits, AM_FMT_SLAB}; }
template <int bits>
inline am_column nogaps(const uint64_t* d, size_t n) { return am_column{(void*)d, n, bits, AM_FMT_BITSTREAM}; }
template <int bits>
inline am_column baseline(const uint64_t* d, size_t n) { return am_column{(void*)d, n, bits, AM_FMT_BASELINE}; }
template <int bits>
inline am_column tight(const void* d, size_t n) { return am_column{(void*)d, n, bits, AM_FMT_TIGHT}; }

// Workspace owned by the caller (device memory, zero-filled once).
struct Workspace {
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

// ---------------------------------------------------------------- binary ops
// src/lib.cuh:4-20 / kernel_binary_op_{baseline,nogaps,slabs}

template <size_t bits, size_t outbits, am_binop op>
inline int binary_op_slabs(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n, void* stream = nullptr)
{
    am_column a = slabs<bits>(d_lhs, n), b = slabs<bits>(d_rhs, n), o = slabs<outbits>(d_out, n);
    return am_binary(op, &a, &b, &o, stream);
}

template <size_t bits, size_t outbits, am_binop op>
inline int binary_op_nogaps(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n, void* stream = nullptr)
{
    am_column a = nogaps<bits>(d_lhs, n), b = nogaps<bits>(d_rhs, n), o = nogaps<outbits>(d_out, n);
    return am_binary(op, &a, &b, &o, stream);
}

template <size_t bits, am_binop op, size_t outbits = 64>
inline int binary_op_baseline(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n, void* stream = nullptr)
{
    am_column a = baseline<bits>(d_lhs, n), b = baseline<bits>(d_rhs, n), o = baseline<outbits>(d_out, n);
    return am_binary(op, &a, &b, &o, stream);
}

// ---------------------------------------------------------------- filters
// src/lib.cuh:43-46 / kernel_filter_{baseline,nogaps,slabs}; the reference
// predicate FILTER_PREDICATE_OP<bits> is v >= 2^(bits-1) (the default here).

template <size_t bits>
constexpr am_predicate reference_predicate() { return am_predicate{AM_PRED_GE, 0, 1ull << (bits - 1)}; }

template <size_t bits>
inline int filter_slabs(const uint64_t* d_in, uint64_t* d_mask, size_t n, am_predicate p = reference_predicate<bits>(),
                        void* stream = nullptr)
{
    am_column c = slabs<bits>(d_in, n);
    return am_filter(&c, p, d_mask, stream);
}
template <size_t bits>
inline int filter_nogaps(const uint64_t* d_in, uint64_t* d_mask, size_t n, am_predicate p = reference_predicate<bits>(),
                         void* stream = nullptr)
{
    am_column c = nogaps<bits>(d_in, n);
    return am_filter(&c, p, d_mask, stream);
}
template <size_t bits>
inline int filter_baseline(const uint64_t* d_in, uint64_t* d_mask, size_t n,
                           am_predicate p = reference_predicate<bits>(), void* stream = nullptr)
{
    am_column c = baseline<bits>(d_in, n);
    return am_filter(&c, p, d_mask, stream);
}

// ---------------------------------------------------------------- compressstore
// kernel_compresstore_slabs_* + the CUB scan (src/kernels/compressstore_slabs.cuh),
// one launch; the mask is the LSB-first bit_iterator byte mask.

template <size_t bits, size_t outbits = bits>
inline int compressstore_slabs(const uint8_t* d_mask, const uint64_t* d_in, uint64_t* d_out, size_t n,
                               uint64_t* d_stats, const Workspace& ws, void* stream = nullptr)
{
    am_column i = slabs<bits>(d_in, n), o = slabs<outbits>(d_out, n);
    return am_compressstore(&i, (const uint64_t*)d_mask, nullptr, &o, nullptr, d_stats, ws.ptr, ws.bytes, stream);
}

// ---------------------------------------------------------------- sum ("hop")
// kernel_hop_slabs / kernel_hop_nogaps with HOP_INIT 0 and HOP_OP + (src/config.cuh:63-71)

template <size_t bits>
inline int hop_slabs(const uint64_t* d_in, size_t n, uint64_t* d_sum, const Workspace& ws, void* stream = nullptr)
{
    am_column c = slabs<bits>(d_in, n);
    return am_sum(&c, d_sum, ws.ptr, ws.bytes, stream);
}
template <size_t bits>
inline int hop_nogaps(const uint64_t* d_in, size_t n, uint64_t* d_sum, const Workspace& ws, void* stream = nullptr)
{
    am_column c = nogaps<bits>(d_in, n);
    return am_sum(&c, d_sum, ws.ptr, ws.bytes, stream);
}

// ---------------------------------------------------------------- group-by
// kernel_groupby_sm_array_{gather,writeout} (src/kernels/groupby_sm_array.cuh);
// default mode keeps the reference's SUM mod 2^bits, key-sorted.

template <size_t bits>
inline int groupby_sm_array(const uint64_t* d_keys, const uint64_t* d_values, size_t n, uint64_t* d_keys_out,
                            uint64_t* d_sums_out, uint64_t* d_result, const Workspace& ws,
                            am_groupby_mode mode = AM_GB_REF, void* stream = nullptr)
{
    am_column k = slabs<bits>(d_keys, n), v = slabs<bits>(d_values, n);
    return am_groupby(&k, &v, nullptr, mode, d_keys_out, d_sums_out, nullptr, d_result, ws.ptr, ws.bytes, stream);
}

} // namespace amethyst
