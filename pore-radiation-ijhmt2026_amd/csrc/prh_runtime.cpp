// Runtime glue of the porerad HIP library: error strings, the library stream,
// grow-only workspaces, PCG64 seeding and LCG jump tables (host side).
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "prh_internal.h"

namespace {
thread_local char g_error[1024] = "";
}

namespace prh {

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int current_stream(hipStream_t *s)
{
    // One non-blocking stream per device at the default priority.
    static std::mutex mu;
    static hipStream_t streams[64] = {};
    int dev = 0;
    PRH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return PRH_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!streams[dev]) PRH_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    *s = streams[dev];
    return PRH_OK;
}

int ensure_capacity(void **buf, size_t *cap, size_t bytes)
{
    if (bytes <= *cap && *buf) return PRH_OK;
    if (*buf) PRH_HIP(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    size_t want = bytes < 256 ? 256 : bytes;
    PRH_HIP(hipMalloc(buf, want));
    *cap = want;
    return PRH_OK;
}

// ------------------------------------------------------------------ PCG64
typedef unsigned __int128 h128;

static h128 h_mult() { return (((h128)PCG_MULT_HI) << 64) | (h128)PCG_MULT_LO; }

void host_jump_tables(JumpTables *t)
{
    // A_b = M^(2^b), G_b = sum_{i < 2^b} M^i (mod 2^128).
    h128 A = h_mult(), G = 1;
    for (int b = 0; b < 64; ++b) {
        t->A[b] = make_u128((uint64_t)(A >> 64), (uint64_t)A);
        t->G[b] = make_u128((uint64_t)(G >> 64), (uint64_t)G);
        G = G * (A + 1);
        A = A * A;
    }
}

// SeedSequence (numpy bit_generator.pyx) followed by PCG64 setseq seeding.
void host_seed_words(const uint32_t *words, int n_words, uint64_t out[4])
{
    const uint32_t INIT_A = 0x43b0d7e5u, MULT_A = 0x931e8875u, INIT_B = 0x8b51f9ddu, MULT_B = 0x58f38dedu;
    const uint32_t MIX_L = 0xca01f9ddu, MIX_R = 0x4973f715u;
    auto hashmix = [&](uint32_t v, uint32_t &hc) {
        v ^= hc;
        hc *= MULT_A;
        v *= hc;
        v ^= v >> 16;
        return v;
    };
    auto mix = [&](uint32_t x, uint32_t y) {
        uint32_t r = MIX_L * x - MIX_R * y;
        r ^= r >> 16;
        return r;
    };
    uint32_t pool[4];
    uint32_t hc = INIT_A;
    for (int i = 0; i < 4; ++i) pool[i] = hashmix(i < n_words ? words[i] : 0u, hc);
    for (int s = 0; s < 4; ++s)
        for (int d = 0; d < 4; ++d)
            if (s != d) pool[d] = mix(pool[d], hashmix(pool[s], hc));
    for (int s = 4; s < n_words; ++s)
        for (int d = 0; d < 4; ++d) pool[d] = mix(pool[d], hashmix(words[s], hc));
    uint32_t w32[8];
    uint32_t hb = INIT_B;
    for (int i = 0; i < 8; ++i) {
        uint32_t v = pool[i & 3];
        v ^= hb;
        hb *= MULT_B;
        v *= hb;
        v ^= v >> 16;
        w32[i] = v;
    }
    uint64_t w[4];
    for (int i = 0; i < 4; ++i) w[i] = (uint64_t)w32[2 * i] | ((uint64_t)w32[2 * i + 1] << 32);
    h128 initstate = (((h128)w[0]) << 64) | w[1];
    h128 initseq = (((h128)w[2]) << 64) | w[3];
    h128 inc = (initseq << 1) | 1u;
    h128 st = 0;
    st = st * h_mult() + inc;
    st += initstate;
    st = st * h_mult() + inc;
    out[0] = (uint64_t)(st >> 64);
    out[1] = (uint64_t)st;
    out[2] = (uint64_t)(inc >> 64);
    out[3] = (uint64_t)inc;
}

}  // namespace prh

extern "C" {

const char *prh_last_error(void) { return g_error; }

const char *prh_version(void) { return "porerad-mi355x 0.1 (gfx950)"; }

int prh_pcg64_seed(const uint32_t *words, int32_t n_words, uint64_t state_inc[4])
{
    PRH_CHECK_ARG(words && n_words > 0 && state_inc, "prh_pcg64_seed: need at least one seed word");
    prh::host_seed_words(words, n_words, state_inc);
    return PRH_OK;
}

}  // extern "C"
