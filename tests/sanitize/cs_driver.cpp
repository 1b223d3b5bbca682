// Host-only driver for the codestream parser (bmi_codestream.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer.  For every file named on
// the command line it opens the stream in both numbps modes and builds the
// code-block table of every tile, then does the same for seeded mutations of
// it: truncation at many lengths, random byte flips, 0xFF insertions (marker
// look-alikes) and rewritten 16/32-bit fields.  Any return code is
// acceptable; a memory or UB error aborts the process (non-zero exit).
// No GPU: the plane addresses are never dereferenced by the parser.
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <vector>

#include "bmi_codestream.h"
#include "bmi_t1_style.h"

int bmi_fail(int code, const char *, ...) { return code; }

extern "C" int BMI_t1_num_segments(int num_pass, int style)
{
    int n = 0;
    for (int i = 0; i < num_pass; ++i) n += bmi_pass_new_segment(i, style) ? 1 : 0;
    return n;
}

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t next()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static long g_opened = 0, g_rejected = 0;

static void exercise(const std::vector<uint8_t> &d)
{
    for (int flags = 0; flags < 2; ++flags) {
        BmiCodestream *cs = nullptr;
        BmiCodestreamInfo info;
        int rc = BMI_codestream_open(d.data(), d.size(), flags, &cs, &info);
        if (rc != BMI_OK) {
            ++g_rejected;
            continue;
        }
        ++g_opened;
        const int nt = info.ntiles_x * info.ntiles_y;
        for (int t = 0; t < nt; ++t) {
            int x0, y0, w, h, nb;
            size_t sb;
            if (BMI_codestream_tile_info(cs, t, &x0, &y0, &w, &h, &nb, &sb) != BMI_OK) continue;
            std::vector<BmiT1Block> blocks((size_t)nb + 1);
            std::vector<uint8_t> segs(sb + 16);
            void *planes[4] = {(void *)0x10000000, (void *)0x20000000, (void *)0x30000000,
                               (void *)0x40000000};
            BmiTileDesc desc;
            BMI_codestream_tile_blocks(cs, t, planes, w, 0, blocks.data(), segs.data(), segs.size(),
                                       &desc);
        }
        BMI_codestream_close(cs);
    }
}

int main(int argc, char **argv)
{
    for (int a = 1; a < argc; ++a) {
        std::ifstream f(argv[a], std::ios::binary);
        std::vector<uint8_t> d((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        if (d.empty()) {
            fprintf(stderr, "cannot read %s\n", argv[a]);
            return 2;
        }
        exercise(d);
        // truncations
        for (size_t n = 0; n < d.size(); n += 1 + d.size() / 97)
            exercise(std::vector<uint8_t>(d.begin(), d.begin() + n));
        // byte flips, 0xFF insertions, rewritten 16 / 32-bit fields
        for (int m = 0; m < 300; ++m) {
            std::vector<uint8_t> e = d;
            const int kind = m % 3;
            const int edits = 1 + (int)(next() % 4);
            for (int k = 0; k < edits && !e.empty(); ++k) {
                const size_t at = next() % e.size();
                if (kind == 0) {
                    e[at] ^= (uint8_t)(1u << (next() % 8));
                } else if (kind == 1) {
                    e.insert(e.begin() + at, 0xFF);
                } else {
                    const uint32_t v = (uint32_t)next();
                    for (size_t b = 0; b < 4 && at + b < e.size(); ++b) e[at + b] = (uint8_t)(v >> (8 * b));
                }
            }
            exercise(e);
        }
    }
    printf("opened %ld, rejected %ld\n", g_opened, g_rejected);
    return 0;
}
