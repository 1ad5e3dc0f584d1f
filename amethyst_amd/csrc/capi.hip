// capi.hip -- helper entry points of the C ABI (sizes, width rules,
// workspace set-up). Operators live with their kernels.
#include <atomic>
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "am_device.hpp"
#include "am_host.hpp"

namespace am {

__global__ __launch_bounds__(256) void k_zero_words(uint64_t* p, uint64_t words)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

} // namespace am

using namespace am;

extern "C" int am_abi_version(void) { return AM_ABI_VERSION; }

extern "C" const char* am_status_string(int status)
{
    switch (status) {
        case AM_OK: return "ok";
        case AM_E_INVALID: return "invalid argument";
        case AM_E_UNSUPPORTED_CONFIG: return "unsupported configuration";
        case AM_E_HIP: return "HIP runtime error";
        case AM_E_WORKSPACE: return "workspace missing or too small";
        case AM_E_CAPACITY: return "output capacity too small";
        case AM_E_ABORTED: return "look-back wait timed out; workspace poisoned until re-initialised";
        case AM_E_SCRATCH_LIMIT: return "input exceeds the fixed scratch of the no-allocation entry point";
    }
    return "unknown status";
}

extern "C" uint64_t am_packed_words(uint32_t format, uint32_t bits, uint64_t n)
{
    if (!valid_format(format, bits)) return 0;
    switch (format) {
        case AM_FMT_BASELINE: return n;
        case AM_FMT_SLAB: {
            uint64_t eps = 64 / bits;
            return (n + eps - 1) / eps;
        }
        case AM_FMT_BITSTREAM: return (n * bits + 63) / 64;
        case AM_FMT_TIGHT: return (n * tight_bytes(bits) + 7) / 8;
        default: return (n + 1) / 2;
    }
}

extern "C" uint32_t am_outbits(uint32_t op, uint32_t bits)
{
    // op_add_outbits / op_mul_outbits / op_mod_outbits, src/config.cuh:44-55
    if (op == AM_OP_ADD) return bits >= 64 ? 64 : bits + 1;
    if (op == AM_OP_MUL) return bits * 2 >= 64 ? 64 : bits * 2;
    return bits;
}

static uint32_t width_of(uint64_t v)
{
    uint32_t w = 1;
    while (w < 64 && (v >> w)) w++;
    return w;
}

extern "C" uint32_t am_pred_width(am_predicate pred, uint32_t bits)
{
    uint64_t vmax = lowmask(bits);
    switch (pred.op) {
        case AM_PRED_LT:
            if (pred.c == 0) return 1;
            vmax = pred.c - 1 < vmax ? pred.c - 1 : vmax;
            break;
        case AM_PRED_LE: vmax = pred.c < vmax ? pred.c : vmax; break;
        case AM_PRED_EQ: vmax = pred.c & vmax; break;
        default: break;
    }
    return width_of(vmax);
}

extern "C" uint64_t am_gen_bitmask(uint32_t bits) { return bits == 0 ? 0 : lowmask(bits); }

extern "C" uint64_t am_workspace_bytes(uint64_t max_elements)
{
    // fixed part (sum partials, group-by tables) + the compaction descriptors
    return WS_FIXED + am_compress_desc_bytes(max_elements) + 64;
}

extern "C" int am_workspace_status(const void* ws)
{
    if (!ws) return AM_E_INVALID;
    uint32_t abort = 0;
    if (hipMemcpy(&abort, (const uint8_t*)ws + offsetof(WsHeader, abort), sizeof(abort), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return AM_E_HIP;
    return abort ? AM_E_ABORTED : AM_OK;
}

extern "C" int am_read_device(void* host_dst, const void* dev_src, uint64_t bytes, void* stream)
{
    if (bytes == 0) return AM_OK;
    if (!host_dst || !dev_src) return AM_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(host_dst, dev_src, bytes, hipMemcpyDeviceToHost, s) != hipSuccess) return AM_E_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? AM_OK : AM_E_HIP;
}

extern "C" int am_workspace_init(void* ws, uint64_t ws_bytes, void* stream)
{
    if (!ws || ws_bytes < WS_FIXED) return AM_E_WORKSPACE;
    uint64_t words = ws_bytes / 8;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero_words, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, (uint64_t*)ws, words);
    return check_launch();
}

// ------------------------------------------------------------ launch-shape table

static std::atomic<uint16_t> g_tune[AM_KNOB_COUNT][65][5][5];

uint32_t am_tuned(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format)
{
    if (knob >= AM_KNOB_COUNT || bits > 64 || in_format > 4 || out_format > 4) return 0;
    return g_tune[knob][bits][in_format][out_format].load(std::memory_order_relaxed);
}

extern "C" int am_tune_set(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format, uint32_t value)
{
    if (knob >= AM_KNOB_COUNT || bits < 1 || bits > 64 || in_format > 4 || out_format > 4 || value > 65535)
        return AM_E_INVALID;
    g_tune[knob][bits][in_format][out_format].store((uint16_t)value, std::memory_order_relaxed);
    return AM_OK;
}

extern "C" uint32_t am_tune_get(uint32_t knob, uint32_t bits, uint32_t in_format, uint32_t out_format)
{
    return am_tuned(knob, bits, in_format, out_format);
}

extern "C" void am_tune_clear(void)
{
    for (auto& k : g_tune)
        for (auto& b : k)
            for (auto& i : b)
                for (auto& o : i) o.store(0, std::memory_order_relaxed);
}

extern "C" int am_tune_load(const char* path)
{
    if (!path) return -1;
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[256];
    int count = 0;
    while (fgets(line, sizeof line, f)) {
        char* h = strchr(line, '#');
        if (h) *h = 0;
        unsigned k, b, fi, fo, v;
        char extra;
        const int got = sscanf(line, "%u %u %u %u %u %c", &k, &b, &fi, &fo, &v, &extra);
        if (got <= 0) continue; // blank or comment
        if (got != 5 || am_tune_set(k, b, fi, fo, v) != AM_OK) {
            fclose(f);
            return -1;
        }
        count++;
    }
    fclose(f);
    return count;
}
