// amethyst.hpp -- header-only C++ facade over the C ABI (amethyst.h) that keeps
// the reference's templated operator surface: compile-time <bits, outbits,
// op> wrappers named after the reference's host wrappers (src/lib.cuh:4-46)
// and kernels (src/kernels/*), the reference's operator functions
// (op_add<bits> ..., src/config.cuh:5-42) as the `op` template argument,
// plus bit_iterator / slab_iterator views (src/bit_iterator.cuh,
// src/slab_iterator.cuh). Every operator call forwards to the gfx950 library;
// nothing here computes on the host except the iterators and the operator
// functions (which the wrappers only compare, at compile time, to find the
// library operator and its result width).
//
// Differences from the reference, all deliberate:
//   * errors come back as am_status codes (the reference aborts via CUDA_TRY);
//   * every wrapper takes a stream (default: the null stream);
//   * block and grid sizes are accepted where the reference takes them and
//     ignored: launch shapes are the library's (its autotuned table);
//   * compressstore_slabs needs a workspace (see Workspace) and reports its
//     count on the device; compressstore_slabs_count() synchronises and
//     returns it. compressstore_baseline_cub returns the count as the
//     reference does and allocates nothing.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <type_traits>
#include <utility>

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

// ---------------------------------------------------------------- operators
// The reference's operator functions (src/config.cuh:5-42): the result is
// masked to `bits`. Division and modulo by zero are defined as in the
// library (x/0 = all ones, x%0 = x; the reference leaves them undefined).
// op_and / op_or / op_xor / op_sub are this library's extensions.

using binop_fn = uint64_t (*)(uint64_t, uint64_t);

template <int bits> constexpr uint64_t op_add(uint64_t l, uint64_t r) { return (l + r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_math(uint64_t l, uint64_t r)
{
    return ((l * l + r * r + 4 * l) / (l * l + r * r + 5 * l)) & gen_bitmask(bits);
}
template <int bits> constexpr uint64_t op_mod(uint64_t l, uint64_t r) { return (r ? l % r : l) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_div(uint64_t l, uint64_t r) { return (r ? l / r : ~0ull) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_mul(uint64_t l, uint64_t r) { return (l * r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_bool_simple(uint64_t l, uint64_t r) { return (l || r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_bool_complex(uint64_t l, uint64_t r)
{
    return ((l && !r) || (!l && r)) & gen_bitmask(bits);
}
template <int bits> constexpr uint64_t op_and(uint64_t l, uint64_t r) { return (l & r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_or(uint64_t l, uint64_t r) { return (l | r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_xor(uint64_t l, uint64_t r) { return (l ^ r) & gen_bitmask(bits); }
template <int bits> constexpr uint64_t op_sub(uint64_t l, uint64_t r) { return (l - r) & gen_bitmask(bits); }

// The library operator and mask width of an operator function (id -1 when
// `op` is none of the functions above). A run-time lookup over their
// addresses (two functions' addresses are not comparable in a constant
// expression); the wrappers return AM_E_INVALID for an unknown operator.
struct op_info {
    int id;
    int bits;
};
template <int... K>
inline op_info find_op(binop_fn op, std::integer_sequence<int, K...>)
{
    op_info r{-1, 0};
    auto chk = [&](binop_fn f, int id, int k) {
        if (r.id < 0 && op == f) r = op_info{id, k};
    };
    ((chk(&op_add<K + 1>, AM_OP_ADD, K + 1), chk(&op_mul<K + 1>, AM_OP_MUL, K + 1),
      chk(&op_mod<K + 1>, AM_OP_MOD, K + 1), chk(&op_div<K + 1>, AM_OP_DIV, K + 1),
      chk(&op_math<K + 1>, AM_OP_MATH, K + 1), chk(&op_bool_simple<K + 1>, AM_OP_BOOL_SIMPLE, K + 1),
      chk(&op_bool_complex<K + 1>, AM_OP_BOOL_COMPLEX, K + 1), chk(&op_and<K + 1>, AM_OP_AND, K + 1),
      chk(&op_or<K + 1>, AM_OP_OR, K + 1), chk(&op_xor<K + 1>, AM_OP_XOR, K + 1),
      chk(&op_sub<K + 1>, AM_OP_SUB, K + 1)),
     ...);
    return r;
}
inline op_info find_op(binop_fn op) { return find_op(op, std::make_integer_sequence<int, 64>{}); }

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

// The reference's form: the operator is one of its operator functions,
// e.g. binary_op_slabs<11, 12, op_add<12>, 1024>(l, r, out, n, grid)
// (src/lib.cuh:17-20). The function's mask width k must not be narrower
// than the output column (the library masks results to the column width;
// k >= outbits gives the same values): otherwise, or for a function that is
// not one of the operator functions above, AM_E_INVALID. block_size and
// grid_size are ignored.
template <size_t bits, size_t outbits, binop_fn op, size_t block_size = 0>
inline int binary_op_slabs(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n,
                           size_t grid_size = 0, void* stream = nullptr)
{
    const op_info f = find_op(op);
    if (f.id < 0 || f.bits < (int)outbits) return AM_E_INVALID;
    (void)grid_size;
    am_column a = slabs<bits>(d_lhs, n), b = slabs<bits>(d_rhs, n), o = slabs<outbits>(d_out, n);
    return am_binary((uint32_t)f.id, &a, &b, &o, stream);
}

template <size_t bits, size_t outbits, binop_fn op, size_t block_size = 0>
inline int binary_op_nogaps(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n,
                            size_t grid_size = 0, void* stream = nullptr)
{
    const op_info f = find_op(op);
    if (f.id < 0 || f.bits < (int)outbits) return AM_E_INVALID;
    (void)grid_size;
    am_column a = nogaps<bits>(d_lhs, n), b = nogaps<bits>(d_rhs, n), o = nogaps<outbits>(d_out, n);
    return am_binary((uint32_t)f.id, &a, &b, &o, stream);
}

// src/lib.cuh:4-7: one uint64 per element in and out; results masked to the
// operator's width k.
template <size_t bits, binop_fn op>
inline int binary_op_baseline(const uint64_t* d_lhs, const uint64_t* d_rhs, uint64_t* d_out, size_t n,
                              void* stream = nullptr)
{
    const op_info f = find_op(op);
    if (f.id < 0) return AM_E_INVALID;
    am_column a = baseline<bits>(d_lhs, n), b = baseline<bits>(d_rhs, n);
    am_column o{(void*)d_out, n, (uint32_t)f.bits, AM_FMT_BASELINE};
    return am_binary((uint32_t)f.id, &a, &b, &o, stream);
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

// The count written to d_stats[0] by compressstore_slabs, after the work
// queued on `stream` (synchronous).
inline size_t compressstore_slabs_count(const uint64_t* d_stats, void* stream = nullptr)
{
    uint64_t k = 0;
    return am_read_device(&k, d_stats, sizeof k, stream) == AM_OK ? (size_t)k : (size_t)-1;
}

// src/lib.cuh:24-39: order-preserving selection of whole uint64 elements by
// an LSB-first byte mask, returning the selected count (synchronous). The
// reference runs cub::DeviceSelect::Flagged with temporary storage and a
// count slot it allocates per call; here the library's no-allocation
// compaction runs with its fixed device scratch (inputs up to
// AM_NOALLOC_MAX_ELEMENTS), and the caller passes the two device words the
// statistics go to (d_stats) -- the one argument call sites add. Returns
// (size_t)-1 on an error.
template <size_t bits>
inline size_t compressstore_baseline_cub(const uint8_t* d_mask, const uint64_t* d_elements, uint64_t* d_out,
                                         size_t n, uint64_t* d_stats, void* stream = nullptr)
{
    am_column i = baseline<64>(d_elements, n), o = baseline<64>(d_out, n);
    if (am_compressstore_noalloc(&i, (const uint64_t*)d_mask, nullptr, &o, nullptr, d_stats, stream) != AM_OK)
        return (size_t)-1;
    return compressstore_slabs_count(d_stats, stream);
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
