// Host sanitizer harness for the DSL reader (uxxxxxxxxh_amd/csrc/dsl.cpp),
// built with g++ -fsanitize=address,undefined by tests/test_host_asan.py.
// For every input file it parses the whole text, every prefix of it, and a
// seeded set of byte mutations, through both readers (single stencil and
// program).  The readers must return a status for each -- an out-of-bounds
// access, an overflow or a leak aborts the run under the sanitizers.
// Prints "<parsed whole> <prefixes ok> <mutations> <accepted mutations>" per
// file.
#include "uhs.h"
#include "uhs_dsl.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// The library's error recorder lives in the HIP part of the host library;
// the reader only needs somewhere to put the message.
static char g_msg[512];
int uhs_set_error(int code, const char *msg)
{
    std::snprintf(g_msg, sizeof g_msg, "%s", msg ? msg : "");
    return code;
}

static int parse_both(const std::string &text)
{
    // Heap copies of exactly the text's size, so a read past its end is seen.
    char *buf = static_cast<char *>(std::malloc(text.size() + 1));
    std::memcpy(buf, text.c_str(), text.size() + 1);
    uhs_dsl_info info;
    const int a = uhs_parse_dsl(buf, &info);
    auto *prog = static_cast<uhs_dsl_program *>(std::malloc(sizeof(uhs_dsl_program)));
    const int b = uhs_parse_dsl_program(buf, prog);
    if (b == UHS_OK) {
        uhs_stencil producers[UHS_MAX_PRODUCERS];
        uhs_program pg;
        (void)uhs_dsl_program_view(prog, producers, &pg);
    }
    std::free(prog);
    std::free(buf);
    return (a == UHS_OK) | ((b == UHS_OK) << 1);
}

int main(int argc, char **argv)
{
    const int mutations = argc > 1 ? std::atoi(argv[1]) : 200;
    for (int f = 2; f < argc; f++) {
        FILE *fp = std::fopen(argv[f], "rb");
        if (!fp) { std::printf("cannot open %s\n", argv[f]); return 2; }
        std::string text;
        char chunk[4096];
        size_t n;
        while ((n = std::fread(chunk, 1, sizeof chunk, fp)) > 0) text.append(chunk, n);
        std::fclose(fp);
        const int whole = parse_both(text);
        int prefixes = 0;
        for (size_t len = 0; len < text.size(); len++) {
            parse_both(text.substr(0, len));
            prefixes++;
        }
        // Mutations: replace, delete or insert one byte, from a fixed seed.
        uint64_t s = 0x9E3779B97F4A7C15ull ^ (uint64_t)f;
        auto next = [&]() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; };
        static const char alphabet[] = "[]{}()+-*#.,0123456789xyzabc \n\t;=";
        int accepted = 0;
        for (int m = 0; m < mutations && !text.empty(); m++) {
            std::string t = text;
            const size_t pos = next() % t.size();
            const char c = alphabet[next() % (sizeof alphabet - 1)];
            switch (next() % 3) {
            case 0: t[pos] = c; break;
            case 1: t.erase(pos, 1); break;
            default: t.insert(pos, 1, c); break;
            }
            accepted += parse_both(t) != 0;
        }
        std::printf("%d %d %d %d\n", whole, prefixes, mutations, accepted);
    }
    return 0;
}
