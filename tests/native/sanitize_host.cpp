// sanitize_host.cpp -- drives the library's host C++ (the GemCAD .asc reader
// and schedule-to-planes builder, the stone-width solver, the camera and cut
// builders) under AddressSanitizer and UndefinedBehaviorSanitizer: the four
// catalogue designs, hand-written malformed inputs and seeded mutations of
// the catalogue bytes (flips, truncations, insertions, splices), every
// accessor on what parses and the width of every closed solid.  Exits 0 when
// every call returned; any sanitizer report aborts the process.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ixg.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint32_t next_u32() {  // xorshift64*
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (uint32_t)((rng_state * 0x2545F4914F6CDD1Dull) >> 32);
}

static int parsed = 0, rejected = 0, solids = 0;

static void exercise(const ixg_asc* s) {
    ixg_asc_info info;
    if (ixg_asc_info_get(s, &info) != IXG_OK) return;
    for (int32_t field = 0; field <= 5; ++field) {
        for (uint32_t i = 0; i < 64; ++i) {
            size_t n = ixg_asc_text(s, field, i, nullptr, 0);
            std::vector<char> buf(n + 1);
            ixg_asc_text(s, field, i, buf.data(), buf.size());
            char small[3];
            ixg_asc_text(s, field, i, small, sizeof small);  // truncated copy
        }
    }
    for (uint32_t t = 0; t < 1024; ++t) {
        ixg_asc_tier tier;
        if (ixg_asc_tier_get(s, t, &tier, nullptr, 0) != IXG_OK && ixg_asc_tier_get(s, t, &tier, nullptr, 0) != IXG_RESOURCE_LIMIT) break;
        std::vector<double> idx(tier.n_indices + 1);
        ixg_asc_tier_get(s, t, &tier, idx.data(), (uint32_t)idx.size());
        for (uint32_t k = 0; k <= tier.n_index_names; ++k) {
            uint32_t pos = 0;
            char name[8];
            ixg_asc_index_name(s, t, k, &pos, name, sizeof name);
        }
    }
    uint32_t n = 0;
    ixg_asc_planes(s, nullptr, 0, &n);
    std::vector<ixg_facet_plane> planes(n + 1);
    if (ixg_asc_planes(s, planes.data(), (uint32_t)planes.size(), &n) == IXG_OK && n > 0) {
        double w = 0.0;
        if (ixg_model_width(planes.data(), n, &w) == IXG_OK) ++solids;
    }
}

static void parse_both(const std::string& text) {
    ixg_asc* s = nullptr;
    ixg_asc_error err;
    if (ixg_asc_parse_text(text.data(), text.size(), &s, &err) == IXG_OK) {
        ++parsed;
        exercise(s);
        ixg_asc_free(s);
    } else {
        ++rejected;
    }
    s = nullptr;
    if (ixg_asc_parse_bytes(reinterpret_cast<const uint8_t*>(text.data()), text.size(), &s, &err) == IXG_OK) {
        exercise(s);
        ixg_asc_free(s);
    }
}

static std::string read_file(const char* path) {
    std::string out;
    FILE* f = std::fopen(path, "rb");
    if (!f) return out;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    std::fclose(f);
    return out;
}

int main(int argc, char** argv) {
    std::vector<std::string> corpus;
    for (int i = 1; i < argc; ++i) corpus.push_back(read_file(argv[i]));
    const char* fixed[] = {
        "", "\n", "GemCad", "GemCad 5.0\n", "GemCad 5.0\ng\n", "GemCad 5.0\ng 96 0\ny 8 y\nI 1.54\na\n",
        "GemCad 5.0\ng 96 0\ny 8 y\nI 1.54\na 41 0.5 0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16\n",
        "GemCad 5.0\ng 2000000000 0\ny 8 y\nI 1.5\na 1 1 0\n", "GemCad 5.0\ng 96 0\ny 0 n\nI 1.54\na 41 0.5 0\n",
        "GemCad 5.0\ng 96 NaN\ny 8 y\nI inf\na nan -inf 0\n", "\x00\x01\x02garbage\xff",
        "\xef\xbb\xbfGemCad 5.0\r\ng 96 0.0\r\ny 8 y\r\nI 1.54\r\na 41 0.5 0 12 n C1 G notes\r\n",
        "GemCad 5.0\rg 96 0\ry 8 y\rI 1.54\ra 41 0.5 0\r", "GemCad 5.0\ng 96 0\ny 8 y\nI 1.54\na 90 1 0 n\na -90 1 0 n\n",
        "GemCad 5.0\ng 96 0\ny 8 y\nI 1.54\na 41 0.5 96 200 -5 1e300 -1e300 4294967296\n",
    };
    for (const char* f : fixed) parse_both(std::string(f));
    parse_both(std::string("GemCad 5.0\ng 96 0\ny 8 y\nI 1.54\n") + std::string(100000, 'a') + "\n");
    for (const std::string& base : corpus) {
        parse_both(base);
        for (int m = 0; m < 400 && !base.empty(); ++m) {
            std::string t = base;
            const int kind = (int)(next_u32() % 5);
            const size_t at = next_u32() % t.size();
            if (kind == 0) {
                for (int k = 0; k < 4; ++k) t[next_u32() % t.size()] ^= (char)(1u << (next_u32() % 8));
            } else if (kind == 1) {
                t.resize(at);
            } else if (kind == 2) {
                t.insert(at, std::string(1 + next_u32() % 16, (char)(next_u32() & 0xFF)));
            } else if (kind == 3) {
                const std::string& other = corpus[next_u32() % corpus.size()];
                t = t.substr(0, at) + other.substr(next_u32() % other.size());
            } else {
                const char* tokens[] = {" nan", " -0", " 1e39", " 720", " 721", "\n", "\r", " ,", " \xc2\xa0", "\xe2\x88\x92"};
                t.insert(at, tokens[next_u32() % 10]);
            }
            parse_both(t);
        }
    }
    // schedule planes from raw tiers, including degenerate ones
    {
        ixg_asc_tier tiers[3] = {{41.0, 0.5, 4, 0}, {90.0, 1.0, 2, 0}, {-43.0, 0.0, 3, 0}};
        double idx[9] = {0, 24, 48, 72, 12, 60, 0, 32, 64};
        ixg_facet_plane out[32];
        uint32_t raw[9], ends[3];
        uint32_t n = 0;
        for (int32_t teeth : {96, -96, 0, 1, 720, 721})
            for (double ref : {0.0, 0.5, -1e300})
                ixg_asc_schedule_planes(tiers, 3, idx, 9, teeth, ref, out, 32, &n, raw, ends);
        ixg_asc_schedule_planes(tiers, 3, idx, 9, 96, 0.0, out, 1, &n, nullptr, nullptr);
        ixg_asc_schedule_planes(tiers, 3, idx, 9, 96, 0.0, nullptr, 0, &n, raw, ends);
        ixg_asc_schedule_planes(tiers, 3, idx, 8, 96, 0.0, out, 32, &n, raw, ends);  // fewer indices than tiers claim
    }
    // cuts, cameras and widths
    ixg_facet_plane srb[IXG_MAX_PLANES];
    const uint32_t n_srb = ixg_cut_standard_round_brilliant(srb);
    double w = 0.0;
    ixg_model_width(srb, n_srb, &w);
    ixg_model_width(srb, 0, &w);
    ixg_model_width(srb, 3, &w);
    for (float crown : {20.0f, 34.5f, 50.0f})
        for (float pav : {38.0f, 40.75f, 45.0f}) {
            ixg_facet_plane rb[IXG_MAX_PLANES];
            const uint32_t n = ixg_cut_round_brilliant(crown, pav, 0.57f, rb);
            if (n > 0 && n <= IXG_MAX_PLANES) ixg_model_width(rb, n, &w);
        }
    ixg_camera cam;
    ixg_camera_new(0.3f, 0.45f, 5.0f, 18.0f, 128, 96, 8, &cam);
    ixg_camera_new(0.0f, 1.5707964f, 2.4f, 42.0f, 1, 1, 1, &cam);
    std::printf("{\"parsed\": %d, \"rejected\": %d, \"solids\": %d, \"srb_width\": %.9g}\n", parsed, rejected,
                solids, w);
    return 0;
}
