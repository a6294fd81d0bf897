// asc_design.cpp -- GemCAD .asc design files and the facet planes they cut.
//
// The reader follows indicatrix-formats/src/asc/parse.rs: physical lines,
// `a` records wrapped over continuation lines reassembled, the g / y / I
// header records required and range-checked, everything else lenient with a
// warning (duplicate gear line, glued keywords, stray tokens, free text after
// the last tier).  Raw file bytes are read as UTF-8 when they are valid UTF-8
// and as Windows-1252 otherwise, with lone CR line ends turned into LF
// (asc/decode.rs, encoding.rs).
//
// The plane builder follows geometry/cuts/asc_schedule.rs: one plane per
// index (one per index-less tier), theta = |angle| in f32 radians, the side
// read from the angle's sign (a sign-negative zero is the culet), azimuth
// 2 pi (index + reference) / |teeth|, offset -|mast|; then near-duplicate
// planes are dropped, keeping the first of each group.
//
// Host-only C++: a design is parsed once and its planes feed the tracer and
// the metrics kernels.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>
#include <vector>

#include "../../include/ixg.h"

namespace {

// ---------------------------------------------------------------- text ---

// Strict UTF-8 validation (no overlong forms, no surrogates, <= U+10FFFF).
bool utf8_valid(const uint8_t* b, size_t n) {
    size_t i = 0;
    while (i < n) {
        uint8_t c = b[i];
        if (c < 0x80) { ++i; continue; }
        size_t len;
        uint32_t cp;
        if (c >= 0xC2 && c <= 0xDF) { len = 2; cp = c & 0x1F; }
        else if (c >= 0xE0 && c <= 0xEF) { len = 3; cp = c & 0x0F; }
        else if (c >= 0xF0 && c <= 0xF4) { len = 4; cp = c & 0x07; }
        else return false;
        if (i + len > n) return false;
        for (size_t k = 1; k < len; ++k) {
            if ((b[i + k] & 0xC0) != 0x80) return false;
            cp = (cp << 6) | (b[i + k] & 0x3F);
        }
        if ((len == 3 && cp < 0x800) || (len == 4 && (cp < 0x10000 || cp > 0x10FFFF))) return false;
        if (cp >= 0xD800 && cp <= 0xDFFF) return false;
        i += len;
    }
    return true;
}

void put_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
        out.push_back((char)cp);
    } else if (cp < 0x800) {
        out.push_back((char)(0xC0 | (cp >> 6)));
        out.push_back((char)(0x80 | (cp & 0x3F)));
    } else if (cp < 0x10000) {
        out.push_back((char)(0xE0 | (cp >> 12)));
        out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
        out.push_back((char)(0x80 | (cp & 0x3F)));
    } else {
        out.push_back((char)(0xF0 | (cp >> 18)));
        out.push_back((char)(0x80 | ((cp >> 12) & 0x3F)));
        out.push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
        out.push_back((char)(0x80 | (cp & 0x3F)));
    }
}

// Windows-1252 bytes 0x80..0x9F (the rest of the code page is Latin-1).
const uint16_t kCp1252High[32] = {
    0x20AC, 0x0081, 0x201A, 0x0192, 0x201E, 0x2026, 0x2020, 0x2021, 0x02C6, 0x2030, 0x0160,
    0x2039, 0x0152, 0x008D, 0x017D, 0x008F, 0x0090, 0x2018, 0x2019, 0x201C, 0x201D, 0x2022,
    0x2013, 0x2014, 0x02DC, 0x2122, 0x0161, 0x203A, 0x0153, 0x009D, 0x017E, 0x0178};

// Decodes one code point of valid UTF-8 at s[i]; returns its byte length.
size_t next_cp(std::string_view s, size_t i, uint32_t* cp) {
    const uint8_t c = (uint8_t)s[i];
    if (c < 0x80) { *cp = c; return 1; }
    size_t len = c >= 0xF0 ? 4 : c >= 0xE0 ? 3 : 2;
    uint32_t v = c & (len == 2 ? 0x1F : len == 3 ? 0x0F : 0x07);
    for (size_t k = 1; k < len && i + k < s.size(); ++k) v = (v << 6) | ((uint8_t)s[i + k] & 0x3F);
    *cp = v;
    return len;
}

// Unicode White_Space (what str::trim / split_whitespace strip).
bool is_space(uint32_t c) {
    return (c >= 0x09 && c <= 0x0D) || c == 0x20 || c == 0x85 || c == 0xA0 || c == 0x1680 ||
           (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F ||
           c == 0x205F || c == 0x3000;
}

std::string_view trim(std::string_view s) {
    size_t a = 0;
    while (a < s.size()) {
        uint32_t cp;
        size_t l = next_cp(s, a, &cp);
        if (!is_space(cp)) break;
        a += l;
    }
    size_t b = a, end = a;
    while (b < s.size()) {
        uint32_t cp;
        size_t l = next_cp(s, b, &cp);
        b += l;
        if (!is_space(cp)) end = b;
    }
    return s.substr(a, end - a);
}

std::vector<std::string_view> split_ws(std::string_view s) {
    std::vector<std::string_view> out;
    size_t i = 0, start = 0;
    bool in_tok = false;
    while (i < s.size()) {
        uint32_t cp;
        size_t l = next_cp(s, i, &cp);
        if (is_space(cp)) {
            if (in_tok) out.push_back(s.substr(start, i - start));
            in_tok = false;
        } else if (!in_tok) {
            in_tok = true;
            start = i;
        }
        i += l;
    }
    if (in_tok) out.push_back(s.substr(start));
    return out;
}

std::string join(const std::vector<std::string_view>& v, size_t from, const char* sep) {
    std::string out;
    for (size_t i = from; i < v.size(); ++i) {
        if (i > from) out += sep;
        out.append(v[i].data(), v[i].size());
    }
    return out;
}

// A string written the way the reference's messages quote one: double quotes
// with backslash escapes for quotes, backslashes and control characters.
std::string quoted(std::string_view s) {
    std::string out = "\"";
    for (size_t i = 0; i < s.size();) {
        uint32_t cp;
        size_t l = next_cp(s, i, &cp);
        if (cp == '"') out += "\\\"";
        else if (cp == '\\') out += "\\\\";
        else if (cp == '\n') out += "\\n";
        else if (cp == '\r') out += "\\r";
        else if (cp == '\t') out += "\\t";
        else if (cp == 0) out += "\\0";
        else if (cp < 0x20 || cp == 0x7F || (cp >= 0x80 && cp < 0xA0)) {
            char buf[16];
            std::snprintf(buf, sizeof buf, "\\u{%x}", cp);
            out += buf;
        } else out.append(s.data() + i, l);
        i += l;
    }
    return out + "\"";
}

// ------------------------------------------------------------- numbers ---

// The decimal grammar of a float literal as the reference reads one: an
// optional sign, then inf / infinity / nan (any case) or digits with an
// optional point and exponent; nothing else (no hex, no surrounding space).
bool parse_plain(std::string_view t, double* out) {
    size_t i = 0;
    bool neg = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) neg = t[i++] == '-';
    std::string_view body = t.substr(i);
    auto ieq = [](std::string_view a, const char* b) {
        size_t n = std::strlen(b);
        if (a.size() != n) return false;
        for (size_t k = 0; k < n; ++k)
            if (std::tolower((unsigned char)a[k]) != b[k]) return false;
        return true;
    };
    if (ieq(body, "inf") || ieq(body, "infinity")) { *out = neg ? -INFINITY : INFINITY; return true; }
    if (ieq(body, "nan")) { *out = NAN; return true; }
    size_t j = 0, digits = 0;
    while (j < body.size() && body[j] >= '0' && body[j] <= '9') { ++j; ++digits; }
    if (j < body.size() && body[j] == '.') {
        ++j;
        while (j < body.size() && body[j] >= '0' && body[j] <= '9') { ++j; ++digits; }
    }
    if (digits == 0) return false;
    if (j < body.size() && (body[j] == 'e' || body[j] == 'E')) {
        ++j;
        if (j < body.size() && (body[j] == '+' || body[j] == '-')) ++j;
        size_t ed = 0;
        while (j < body.size() && body[j] >= '0' && body[j] <= '9') { ++j; ++ed; }
        if (ed == 0) return false;
    }
    if (j != body.size()) return false;
    std::string s(t);
    *out = std::strtod(s.c_str(), nullptr);  // correctly rounded; overflow gives +-inf
    return true;
}

// parse_number (asc/parse.rs:394-400): also reads the typeset minus U+2212.
bool parse_number(std::string_view t, double* out) {
    if (t.find("\xE2\x88\x92") == std::string_view::npos) return parse_plain(t, out);
    std::string s;
    for (size_t i = 0; i < t.size();) {
        if (t.compare(i, 3, "\xE2\x88\x92") == 0) { s.push_back('-'); i += 3; }
        else s.push_back(t[i++]);
    }
    return parse_plain(s, out);
}

// The shortest decimal that reads back as v, written without an exponent
// (how the reference prints a float in a message).
std::string fmt_f64(double v) {
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    std::string sign = std::signbit(v) ? "-" : "";
    double a = std::fabs(v);
    if (a == 0.0) return sign + "0";
    char buf[64];
    int p = 1;
    for (; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, a);
        if (std::strtod(buf, nullptr) == a) break;
    }
    std::string m;  // mantissa digits
    for (const char* c = buf; *c && *c != 'e'; ++c)
        if (*c >= '0' && *c <= '9') m.push_back(*c);
    int e = std::atoi(std::strchr(buf, 'e') + 1);
    while (m.size() > 1 && m.back() == '0') m.pop_back();
    const int n = (int)m.size();
    std::string out;
    if (e >= n - 1) out = m + std::string((size_t)(e - (n - 1)), '0');
    else if (e >= 0) out = m.substr(0, (size_t)e + 1) + "." + m.substr((size_t)e + 1);
    else out = "0." + std::string((size_t)(-e - 1), '0') + m;
    return sign + out;
}

// ------------------------------------------------------------ schedule ---

struct Tier {
    double angle_deg = 0.0, mast = 0.0;
    std::string name, notes;
    std::vector<double> indices;
    std::vector<std::pair<uint32_t, std::string>> index_names;
};

}  // namespace

struct ixg_asc {
    std::string version;
    int32_t gear_teeth = 0;
    double gear_reference_angle = 0.0;
    uint32_t symmetry_order = 0;
    bool mirror = false;
    double refractive_index = 0.0;
    std::vector<std::string> headers, footnotes, warnings;
    std::vector<Tier> tiers;
    int32_t line_ending = IXG_ASC_CRLF;
    uint32_t gear_line = 0;
};

namespace {

struct AscError {
    int32_t kind;
    uint32_t line = 0;
    std::string field, token, text;
    double value = 0.0;
    int64_t teeth = 0;
    uint32_t field_count = 0;
};

[[noreturn]] void fail(int32_t kind, uint32_t line) {
    AscError e;
    e.kind = kind;
    e.line = line;
    throw e;
}
[[noreturn]] void fail_token(int32_t kind, uint32_t line, std::string_view token) {
    AscError e;
    e.kind = kind;
    e.line = line;
    e.token = std::string(token);
    throw e;
}
[[noreturn]] void fail_value(int32_t kind, uint32_t line, const char* field, std::string_view token,
                             double value) {
    AscError e;
    e.kind = kind;
    e.line = line;
    e.field = field;
    e.token = std::string(token);
    e.value = value;
    throw e;
}

// The reference's wording for each error (asc/error.rs Display).
std::string message(const AscError& e) {
    const std::string L = "line " + std::to_string(e.line) + ": ";
    switch (e.kind) {
    case IXG_ASC_ERR_EMPTY_INPUT: return "empty input";
    case IXG_ASC_ERR_INVALID_UTF8: return "text is not valid UTF-8";
    case IXG_ASC_ERR_GEAR_LINE_MISSING_FIELDS:
        return L + "'g' (gear) line needs a tooth count and a reference angle, got " + quoted(e.text);
    case IXG_ASC_ERR_GEAR_TEETH_NOT_NUMERIC: return L + "gear tooth count " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_GEAR_REFERENCE_ANGLE_NOT_NUMERIC:
        return L + "gear reference angle " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_SYMMETRY_LINE_MISSING_FIELDS:
        return L + "'y' (symmetry) line needs an order and a mirror flag, got " + quoted(e.text);
    case IXG_ASC_ERR_SYMMETRY_ORDER_NOT_NUMERIC: return L + "symmetry order " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_REFRACTIVE_INDEX_MISSING: return L + "'I' (refractive index) line has no value";
    case IXG_ASC_ERR_REFRACTIVE_INDEX_NOT_NUMERIC:
        return L + "refractive index " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_NOT_WHOLE_NUMBER:
        return L + e.field + " " + quoted(e.token) + " must be a whole number in range, got " + fmt_f64(e.value);
    case IXG_ASC_ERR_GEAR_TEETH_ZERO: return L + "'g' (gear) tooth count must not be zero";
    case IXG_ASC_ERR_GEAR_TEETH_TOO_LARGE:
        return L + "'g' (gear) tooth count " + std::to_string(e.teeth) + " exceeds the limit of " +
               std::to_string(IXG_ASC_MAX_GEAR_TEETH) + " teeth";
    case IXG_ASC_ERR_MIRROR_FLAG_INVALID:
        return L + "'y' (symmetry) line's mirror flag " + quoted(e.token) + " is neither 'y' nor 'n'";
    case IXG_ASC_ERR_MISSING_GEAR_LINE: return "missing required 'g' (gear teeth) header line";
    case IXG_ASC_ERR_MISSING_SYMMETRY_LINE: return "missing required 'y' (symmetry) header line";
    case IXG_ASC_ERR_MISSING_REFRACTIVE_INDEX_LINE: return "missing required 'I' (refractive index) header line";
    case IXG_ASC_ERR_NO_TIER_RECORDS: return "no valid 'a' (facet tier) records found";
    case IXG_ASC_ERR_TIER_RECORD_TOO_SHORT:
        return L + "'a' record needs at least an angle and a mast distance, got " + std::to_string(e.field_count) +
               " field(s)";
    case IXG_ASC_ERR_ANGLE_NOT_NUMERIC: return L + "angle " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_MAST_NOT_NUMERIC: return L + "mast distance " + quoted(e.token) + " is not numeric";
    case IXG_ASC_ERR_NON_FINITE_VALUE:
        return L + e.field + " " + quoted(e.token) + " is not finite (" + fmt_f64(e.value) + ")";
    case IXG_ASC_ERR_SYMMETRY_ORDER_ZERO: return L + "'y' (symmetry) order must not be zero";
    case IXG_ASC_ERR_REFRACTIVE_INDEX_OUT_OF_RANGE:
        return L + "refractive index " + fmt_f64(e.value) + " must be greater than 1.0";
    default: return "unknown error";
    }
}

constexpr double kIntegralTolerance = 1e-9;  // asc/parse.rs:405

// require_integral_i32 + require_nonzero_gear_teeth (asc/parse.rs:416-481).
int32_t gear_teeth_value(double v, uint32_t line, std::string_view token) {
    if (!std::isfinite(v)) fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, line, "gear tooth count", token, v);
    if (std::fabs(v - std::round(v)) > kIntegralTolerance || std::fabs(v) > 2147483647.0)
        fail_value(IXG_ASC_ERR_NOT_WHOLE_NUMBER, line, "gear tooth count", token, v);
    const int32_t t = (int32_t)std::round(v);
    if (t == 0) fail(IXG_ASC_ERR_GEAR_TEETH_ZERO, line);
    if (std::llabs((long long)t) > IXG_ASC_MAX_GEAR_TEETH) {
        AscError e;
        e.kind = IXG_ASC_ERR_GEAR_TEETH_TOO_LARGE;
        e.line = line;
        e.teeth = t;
        throw e;
    }
    return t;
}

// split_glued_keyword (asc/parse.rs:340-380): "g96 0.0" -> "g 96 0.0",
// "y8y" -> "y 8 y", "I1.54" -> "I 1.54"; empty when the line has no such token.
std::string split_glued(std::string_view line) {
    std::vector<std::string_view> tok = split_ws(line);
    if (tok.empty()) return {};
    std::string_view first = tok[0];
    const char k = first[0];
    if (k != 'g' && k != 'y' && k != 'I') return {};
    std::string_view value = first.substr(1);
    if (value.empty()) return {};
    std::string v;
    double dummy;
    if (k == 'y') {
        char flag = value.back();
        std::string_view order = value;
        bool has_flag = flag == 'y' || flag == 'Y' || flag == 'n' || flag == 'N';
        if (has_flag) order = value.substr(0, value.size() - 1);
        if (!parse_number(order, &dummy)) return {};
        v = std::string(order);
        if (has_flag) v += std::string(" ") + flag;
    } else {
        if (!parse_number(value, &dummy)) return {};
        v = std::string(value);
    }
    std::string_view rest = trim(line.substr(first.size()));
    std::string out = std::string(1, k) + " " + v;
    if (!rest.empty()) out += " " + std::string(rest);
    return out;
}

class Reader {
public:
    explicit Reader(ixg_asc* s) : s_(s) {}

    void line(uint32_t no, std::string_view raw) {
        std::string_view line = trim(raw);
        if (line.empty() || line[0] == ';' || line[0] == '#') return;
        std::string split = split_glued(line);
        if (!split.empty()) {
            warn(no, "keyword glued to its value, read as " + quoted(split));
            line = split;
        }
        std::vector<std::string_view> tok = split_ws(line);
        if (tok.empty()) return;
        const std::string_view first = tok[0];
        if (first == "GemCad" || first == "GemCAD" || first == "Gemcad") {
            flush();
            s_->version = join(tok, 1, " ");
        } else if (first == "g") {
            flush();
            gear(no, line, tok);
        } else if (first == "y") {
            flush();
            symmetry(no, line, tok);
        } else if (first == "I") {
            flush();
            index_of_refraction(no, tok);
        } else if (first == "H") {
            flush();
            s_->headers.emplace_back(trim(line.substr(1)));
        } else if (first == "F") {
            flush();
            s_->footnotes.emplace_back(trim(line.substr(1)));
        } else if (first == "a") {
            flush();
            seen_tier_ = true;
            open_ = true;
            open_line_ = no;
            pending_.assign(tok.begin() + 1, tok.end());  // copies: `line` may be a local split
        } else {
            other(no, line, tok);
        }
    }

    void finish() {
        flush();
        if (!have_gear_) fail(IXG_ASC_ERR_MISSING_GEAR_LINE, 0);
        if (!have_order_) fail(IXG_ASC_ERR_MISSING_SYMMETRY_LINE, 0);
        if (!have_ri_) fail(IXG_ASC_ERR_MISSING_REFRACTIVE_INDEX_LINE, 0);
        if (s_->tiers.empty()) fail(IXG_ASC_ERR_NO_TIER_RECORDS, 0);
    }

private:
    void warn(uint32_t no, const std::string& what) {
        s_->warnings.push_back("line " + std::to_string(no) + ": " + what);
    }

    void gear(uint32_t no, std::string_view line, const std::vector<std::string_view>& tok) {
        if (tok.size() < 3) {
            AscError e;
            e.kind = IXG_ASC_ERR_GEAR_LINE_MISSING_FIELDS;
            e.line = no;
            e.text = std::string(line);
            throw e;
        }
        double teeth, ref;
        if (!parse_plain(tok[1], &teeth)) fail_token(IXG_ASC_ERR_GEAR_TEETH_NOT_NUMERIC, no, tok[1]);
        const int32_t t = gear_teeth_value(teeth, no, tok[1]);
        if (!parse_plain(tok[2], &ref)) fail_token(IXG_ASC_ERR_GEAR_REFERENCE_ANGLE_NOT_NUMERIC, no, tok[2]);
        if (!std::isfinite(ref))
            fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "gear reference angle", tok[2], ref);
        if (have_gear_) warn(no, "duplicate 'g' (gear) line overrides the previous value");
        set_gear(no, t, ref);
    }

    void set_gear(uint32_t no, int32_t teeth, double ref) {
        have_gear_ = true;
        s_->gear_teeth = teeth;
        s_->gear_line = no;
        s_->gear_reference_angle = ref;
    }

    void symmetry(uint32_t no, std::string_view line, const std::vector<std::string_view>& tok) {
        if (tok.size() < 3) {
            AscError e;
            e.kind = IXG_ASC_ERR_SYMMETRY_LINE_MISSING_FIELDS;
            e.line = no;
            e.text = std::string(line);
            throw e;
        }
        double v;
        if (!parse_plain(tok[1], &v)) fail_token(IXG_ASC_ERR_SYMMETRY_ORDER_NOT_NUMERIC, no, tok[1]);
        if (!std::isfinite(v)) fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "symmetry order", tok[1], v);
        if (std::fabs(v - std::round(v)) > kIntegralTolerance || !(v >= 0.0 && v <= 4294967295.0))
            fail_value(IXG_ASC_ERR_NOT_WHOLE_NUMBER, no, "symmetry order", tok[1], v);
        const uint32_t order = (uint32_t)std::round(v);
        if (order == 0) fail(IXG_ASC_ERR_SYMMETRY_ORDER_ZERO, no);
        have_order_ = true;
        s_->symmetry_order = order;
        const std::string_view f = tok[2];
        if (f == "y" || f == "Y") s_->mirror = true;
        else if (f == "n" || f == "N") s_->mirror = false;
        else fail_token(IXG_ASC_ERR_MIRROR_FLAG_INVALID, no, f);
    }

    void index_of_refraction(uint32_t no, const std::vector<std::string_view>& tok) {
        if (tok.size() < 2) fail(IXG_ASC_ERR_REFRACTIVE_INDEX_MISSING, no);
        double ri;
        if (!parse_plain(tok[1], &ri)) fail_token(IXG_ASC_ERR_REFRACTIVE_INDEX_NOT_NUMERIC, no, tok[1]);
        if (!std::isfinite(ri)) fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "refractive index", tok[1], ri);
        if (ri <= 1.0) {
            AscError e;
            e.kind = IXG_ASC_ERR_REFRACTIVE_INDEX_OUT_OF_RANGE;
            e.line = no;
            e.value = ri;
            throw e;
        }
        have_ri_ = true;
        s_->refractive_index = ri;
    }

    // read_other_line (asc/parse.rs:280-320): a continuation of the open
    // tier, the bare "96 0.0" gear line, or stray text.
    void other(uint32_t no, std::string_view line, const std::vector<std::string_view>& tok) {
        if (open_) {
            double d;
            if (parse_plain(tok[0], &d) || tok[0] == "n" || tok[0] == "G")
                pending_.insert(pending_.end(), tok.begin(), tok.end());
            else
                warn(no, "ignored free text after the last facet tier: " + quoted(line));
            return;
        }
        if (!have_gear_ && !seen_tier_) {
            double teeth, ref;
            if (tok.size() == 2 && parse_plain(tok[0], &teeth) && parse_plain(tok[1], &ref) && std::isfinite(ref)) {
                try {
                    const int32_t t = gear_teeth_value(teeth, no, tok[0]);
                    set_gear(no, t, ref);
                    return;
                } catch (const AscError&) {
                }
            }
        }
        warn(no, "ignored unrecognized line: " + quoted(line));
    }

    void flush() {
        if (!open_) return;
        open_ = false;
        s_->tiers.push_back(tier(open_line_, pending_));
        pending_.clear();
    }

    // parse_tier + parse_angle_and_mast (asc/parse.rs:528-678).
    Tier tier(uint32_t no, const std::vector<std::string>& tok) {
        if (tok.size() < 2) {
            AscError e;
            e.kind = IXG_ASC_ERR_TIER_RECORD_TOO_SHORT;
            e.line = no;
            e.field_count = (uint32_t)tok.size();
            throw e;
        }
        Tier t;
        if (!parse_number(tok[0], &t.angle_deg)) fail_token(IXG_ASC_ERR_ANGLE_NOT_NUMERIC, no, tok[0]);
        if (!std::isfinite(t.angle_deg))
            fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "angle", tok[0], t.angle_deg);
        if (!parse_number(tok[1], &t.mast)) fail_token(IXG_ASC_ERR_MAST_NOT_NUMERIC, no, tok[1]);
        if (!std::isfinite(t.mast)) fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "mast distance", tok[1], t.mast);
        if (t.mast < 0.0) {
            if (t.angle_deg == 0.0) {  // the culet: a sign-negative zero angle, positive mast
                t.angle_deg = -0.0;
                t.mast = -t.mast;
            } else {
                warn(no, "negative mast " + fmt_f64(t.mast) + " on a nonzero-angle tier (" + fmt_f64(t.angle_deg) +
                             " deg); its magnitude is used");
            }
        }
        std::vector<std::string> names;
        std::vector<std::string_view> notes;
        bool in_notes = false, expect_name = false;
        for (size_t i = 2; i < tok.size(); ++i) {
            const std::string_view w = tok[i];
            if (in_notes) { notes.push_back(w); continue; }
            if (expect_name) {  // the token after "n" is a name even when it reads as a number or "G"
                expect_name = false;
                const uint32_t at = t.indices.empty() ? 0u : (uint32_t)t.indices.size() - 1u;
                t.index_names.emplace_back(at, std::string(w));
                if (names.empty() || names.back() != w) names.emplace_back(w);
                continue;
            }
            if (w == "n") { expect_name = true; continue; }
            if (w == "G") { in_notes = true; continue; }
            double v;
            if (parse_number(w, &v)) {
                if (!std::isfinite(v)) fail_value(IXG_ASC_ERR_NON_FINITE_VALUE, no, "index", w, v);
                t.indices.push_back(v);
            } else {
                warn(no, "ignored token that is neither a number nor a name/notes marker: " + quoted(w));
            }
        }
        for (size_t i = 0; i < names.size(); ++i) t.name += (i ? "/" : "") + names[i];
        t.notes = join(notes, 0, " ");
        return t;
    }

    ixg_asc* s_;
    bool have_gear_ = false, have_order_ = false, have_ri_ = false;
    bool seen_tier_ = false, open_ = false;
    uint32_t open_line_ = 0;
    std::vector<std::string> pending_;
};

// Lines split on LF (a CR right before it belongs to the line ending); the
// views point into `text`, which outlives the reader.
void parse_text(const std::string& text_in, ixg_asc* s) {
    std::string_view text = text_in;
    if (text.compare(0, 3, "\xEF\xBB\xBF") == 0) text.remove_prefix(3);
    if (trim(text).empty()) fail(IXG_ASC_ERR_EMPTY_INPUT, 0);
    s->line_ending = text.find("\r\n") != std::string_view::npos ? IXG_ASC_CRLF : IXG_ASC_LF;
    Reader r(s);
    uint32_t no = 0;
    size_t at = 0;
    while (at < text.size()) {
        size_t nl = text.find('\n', at);
        size_t end = nl == std::string_view::npos ? text.size() : nl;
        std::string_view ln = text.substr(at, end - at);
        if (!ln.empty() && ln.back() == '\r') ln.remove_suffix(1);
        r.line(++no, ln);
        at = nl == std::string_view::npos ? text.size() : nl + 1;
    }
    r.finish();
}

// decode_asc_bytes (asc/decode.rs): BOM dropped, UTF-8 kept, anything else
// read as Windows-1252; lone CRs become LFs.
std::string decode_bytes(const uint8_t* b, size_t n) {
    if (n >= 3 && b[0] == 0xEF && b[1] == 0xBB && b[2] == 0xBF) { b += 3; n -= 3; }
    std::string text;
    if (utf8_valid(b, n)) {
        text.assign((const char*)b, n);
    } else {
        text.reserve(n + n / 4);
        for (size_t i = 0; i < n; ++i) {
            const uint8_t c = b[i];
            put_utf8(text, (c >= 0x80 && c <= 0x9F) ? kCp1252High[c - 0x80] : c);
        }
    }
    for (size_t i = 0; i < text.size(); ++i)
        if (text[i] == '\r' && (i + 1 == text.size() || text[i + 1] != '\n')) text[i] = '\n';
    return text;
}

void report(const AscError& e, ixg_asc_error* err) {
    if (!err) return;
    std::memset(err, 0, sizeof *err);
    err->kind = e.kind;
    err->line = e.line;
    err->value = e.value;
    err->teeth = (int32_t)e.teeth;
    err->field_count = e.field_count;
    std::snprintf(err->field, sizeof err->field, "%s", e.field.c_str());
    std::snprintf(err->token, sizeof err->token, "%s", e.token.c_str());
    std::snprintf(err->message, sizeof err->message, "%s", message(e).c_str());
}

ixg_status parse_into(const std::string& text, ixg_asc** out, ixg_asc_error* err) {
    std::unique_ptr<ixg_asc> s(new ixg_asc());
    try {
        parse_text(text, s.get());
    } catch (const AscError& e) {
        report(e, err);
        return IXG_INVALID;
    }
    if (err) std::memset(err, 0, sizeof *err);
    *out = s.release();
    return IXG_OK;
}

// No C++ exception crosses the C ABI: an allocation failure (a huge or
// hostile file) is a resource limit, anything else unexpected an invalid
// input, and the caller's process keeps running.
template <class F>
ixg_status guarded(F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return IXG_RESOURCE_LIMIT;
    } catch (const std::length_error&) {
        return IXG_RESOURCE_LIMIT;
    } catch (...) {
        return IXG_INVALID;
    }
}

size_t copy_text(const std::string& s, char* buf, size_t cap) {
    if (buf && cap > 0) {
        size_t n = s.size() < cap - 1 ? s.size() : cap - 1;
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return s.size();
}

// -------------------------------------------------------------- planes ---

constexpr float kPi = 3.14159265358979323846f;

// Correctly rounded f32 sine / cosine, as the other cut builders use.
float sin_cr(float x) { return (float)std::sin((double)x); }
float cos_cr(float x) { return (float)std::cos((double)x); }

ixg_facet_plane make_plane(float x, float y, float z, float d) {
    const float r = 1.0f / std::sqrt((x * x + y * y) + z * z);
    ixg_facet_plane p;
    p.normal[0] = x * r;
    p.normal[1] = y * r;
    p.normal[2] = z * r;
    p.d = d;
    return p;
}

// normals_coincide (cuts/asc_schedule.rs:335-339), in f64.
bool normals_coincide(const ixg_facet_plane& a, const ixg_facet_plane& b) {
    double u[3], v[3];
    for (int k = 0; k < 3; ++k) { u[k] = a.normal[k]; v[k] = b.normal[k]; }
    auto unit = [](double* w) {
        const double r = 1.0 / std::sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
        const bool ok = std::isfinite(r) && r > 0.0;
        for (int k = 0; k < 3; ++k) w[k] = ok ? w[k] * r : 0.0;
    };
    unit(u);
    unit(v);
    return 1.0 - ((u[0] * v[0] + u[1] * v[1]) + u[2] * v[2]) <= 2e-7;
}

}  // namespace

extern "C" {

ixg_status ixg_asc_parse_text(const char* text, size_t len, ixg_asc** out, ixg_asc_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (!out || (!text && len)) return IXG_INVALID;
    *out = nullptr;
    if (!utf8_valid((const uint8_t*)text, len)) {
        AscError e;
        e.kind = IXG_ASC_ERR_INVALID_UTF8;
        report(e, err);
        return IXG_INVALID;
    }
    return guarded([&] { return parse_into(std::string(text ? text : "", len), out, err); });
}

ixg_status ixg_asc_parse_bytes(const uint8_t* bytes, size_t len, ixg_asc** out, ixg_asc_error* err) {
    if (err) std::memset(err, 0, sizeof *err);
    if (!out || (!bytes && len)) return IXG_INVALID;
    *out = nullptr;
    return guarded([&] { return parse_into(decode_bytes(bytes, len), out, err); });
}

void ixg_asc_free(ixg_asc* s) { delete s; }

ixg_status ixg_asc_info_get(const ixg_asc* s, ixg_asc_info* info) {
    if (!s || !info) return IXG_INVALID;
    std::memset(info, 0, sizeof *info);
    info->gear_teeth = s->gear_teeth;
    info->symmetry_order = s->symmetry_order;
    info->mirror = s->mirror ? 1 : 0;
    info->line_ending = s->line_ending;
    info->gear_reference_angle = s->gear_reference_angle;
    info->refractive_index = s->refractive_index;
    info->n_tiers = (uint32_t)s->tiers.size();
    info->n_headers = (uint32_t)s->headers.size();
    info->n_footnotes = (uint32_t)s->footnotes.size();
    info->n_warnings = (uint32_t)s->warnings.size();
    info->gear_line = s->gear_line;
    uint32_t planes = 0;
    for (const Tier& t : s->tiers) planes += t.indices.empty() ? 1u : (uint32_t)t.indices.size();
    info->facet_plane_count = planes;
    return IXG_OK;
}

ixg_status ixg_asc_tier_get(const ixg_asc* s, uint32_t i, ixg_asc_tier* tier, double* indices, uint32_t cap) {
    if (!s || !tier || i >= s->tiers.size()) return IXG_INVALID;
    const Tier& t = s->tiers[i];
    tier->angle_deg = t.angle_deg;
    tier->mast = t.mast;
    tier->n_indices = (uint32_t)t.indices.size();
    tier->n_index_names = (uint32_t)t.index_names.size();
    if (indices) {
        if (cap < t.indices.size()) return IXG_RESOURCE_LIMIT;
        for (size_t k = 0; k < t.indices.size(); ++k) indices[k] = t.indices[k];
    }
    return IXG_OK;
}

size_t ixg_asc_text(const ixg_asc* s, int32_t what, uint32_t i, char* buf, size_t cap) {
    if (buf && cap) buf[0] = '\0';
    if (!s) return 0;
    auto pick = [&](const std::vector<std::string>& v) { return i < v.size() ? copy_text(v[i], buf, cap) : 0; };
    switch (what) {
    case IXG_ASC_TEXT_VERSION: return copy_text(s->version, buf, cap);
    case IXG_ASC_TEXT_HEADER: return pick(s->headers);
    case IXG_ASC_TEXT_FOOTNOTE: return pick(s->footnotes);
    case IXG_ASC_TEXT_WARNING: return pick(s->warnings);
    case IXG_ASC_TEXT_TIER_NAME: return i < s->tiers.size() ? copy_text(s->tiers[i].name, buf, cap) : 0;
    case IXG_ASC_TEXT_TIER_NOTES: return i < s->tiers.size() ? copy_text(s->tiers[i].notes, buf, cap) : 0;
    default: return 0;
    }
}

size_t ixg_asc_index_name(const ixg_asc* s, uint32_t tier, uint32_t k, uint32_t* position, char* buf, size_t cap) {
    if (buf && cap) buf[0] = '\0';
    if (!s || tier >= s->tiers.size() || k >= s->tiers[tier].index_names.size()) return 0;
    const auto& e = s->tiers[tier].index_names[k];
    if (position) *position = e.first;
    return copy_text(e.second, buf, cap);
}

}  // extern "C"

namespace {

ixg_status schedule_planes(const ixg_asc_tier* tiers, uint32_t n_tiers, const double* indices,
                           uint32_t n_indices, int32_t gear_teeth, double gear_reference_angle,
                           ixg_facet_plane* out, uint32_t cap, uint32_t* n_out, uint32_t* raw_to_plane,
                           uint32_t* tier_ends) {
    if (n_out) *n_out = 0;
    if (!n_out || (n_tiers && !tiers)) return IXG_INVALID;
    uint64_t total = 0;
    for (uint32_t t = 0; t < n_tiers; ++t) total += tiers[t].n_indices;
    if (total != n_indices || (n_indices && !indices)) return IXG_INVALID;

    // raw_asc_planes (cuts/asc_schedule.rs:152-200)
    const uint32_t abs_teeth = gear_teeth < 0 ? 0u - (uint32_t)gear_teeth : (uint32_t)gear_teeth;
    const float teeth = (float)(abs_teeth > 1u ? abs_teeth : 1u);
    const float ref = (float)gear_reference_angle;
    std::vector<ixg_facet_plane> raw;
    std::vector<uint32_t> ends;
    raw.reserve(n_indices + n_tiers);
    const double* idx = indices;
    for (uint32_t t = 0; t < n_tiers; ++t) {
        const double a = tiers[t].angle_deg;
        const bool crown = a > 0.0 || (a == 0.0 && !std::signbit(a));  // tier_is_crown_side (plane.rs:25-27)
        const float theta = (float)std::fabs(a) * (kPi / 180.0f);
        const float st = sin_cr(theta), ct = cos_cr(theta);
        const float sy = crown ? ct : -ct;
        const float d = -(float)std::fabs(tiers[t].mast);
        if (tiers[t].n_indices == 0) {
            raw.push_back(make_plane(0.0f, sy, st, d));
        } else {
            for (uint32_t k = 0; k < tiers[t].n_indices; ++k) {
                // index_to_azimuth (cuts/round_brilliant.rs:64-66)
                const float phi = 2.0f * kPi * ((float)idx[k] + ref) / teeth;
                raw.push_back(make_plane(st * cos_cr(phi), sy, st * sin_cr(phi), d));
            }
            idx += tiers[t].n_indices;
        }
        ends.push_back((uint32_t)raw.size());
    }

    // dedup_representatives (cuts/asc_schedule.rs:298-322)
    const float kOffsetRel = 1.0f / 2048.0f;
    std::vector<uint32_t> kept, rep(raw.size());
    for (uint32_t i = 0; i < raw.size(); ++i) {
        uint32_t r = i;
        for (uint32_t k : kept) {
            const float scale = std::fmax(std::fmax(std::fabs(raw[k].d), std::fabs(raw[i].d)), 1.0f);
            if (normals_coincide(raw[k], raw[i]) && std::fabs(raw[k].d - raw[i].d) <= kOffsetRel * scale) {
                r = k;
                break;
            }
        }
        if (r == i) kept.push_back(i);
        rep[i] = r;
    }
    *n_out = (uint32_t)kept.size();
    std::vector<uint32_t> out_index(raw.size()), kept_before(raw.size() + 1, 0);
    uint32_t count = 0;
    for (uint32_t i = 0; i < raw.size(); ++i) {
        out_index[i] = rep[i] == i ? count++ : out_index[rep[i]];
        kept_before[i + 1] = count;
    }
    if (raw_to_plane)
        for (uint32_t i = 0; i < raw.size(); ++i) raw_to_plane[i] = out_index[i];
    if (tier_ends)
        for (uint32_t t = 0; t < n_tiers; ++t) tier_ends[t] = kept_before[ends[t]];
    if (!out) return IXG_OK;
    if (cap < kept.size()) return IXG_RESOURCE_LIMIT;
    for (uint32_t i = 0; i < kept.size(); ++i) out[i] = raw[kept[i]];
    return IXG_OK;
}

ixg_status schedule_planes_of(const ixg_asc* s, ixg_facet_plane* out, uint32_t cap, uint32_t* n_out) {
    std::vector<ixg_asc_tier> tiers(s->tiers.size());
    std::vector<double> indices;
    for (size_t t = 0; t < s->tiers.size(); ++t) {
        tiers[t].angle_deg = s->tiers[t].angle_deg;
        tiers[t].mast = s->tiers[t].mast;
        tiers[t].n_indices = (uint32_t)s->tiers[t].indices.size();
        tiers[t].n_index_names = 0;
        indices.insert(indices.end(), s->tiers[t].indices.begin(), s->tiers[t].indices.end());
    }
    return schedule_planes(tiers.data(), (uint32_t)tiers.size(), indices.data(), (uint32_t)indices.size(),
                           s->gear_teeth, s->gear_reference_angle, out, cap, n_out, nullptr, nullptr);
}

}  // namespace

extern "C" {

ixg_status ixg_asc_schedule_planes(const ixg_asc_tier* tiers, uint32_t n_tiers, const double* indices,
                                   uint32_t n_indices, int32_t gear_teeth, double gear_reference_angle,
                                   ixg_facet_plane* out, uint32_t cap, uint32_t* n_out, uint32_t* raw_to_plane,
                                   uint32_t* tier_ends) {
    if (n_out) *n_out = 0;
    return guarded([&] {
        return schedule_planes(tiers, n_tiers, indices, n_indices, gear_teeth, gear_reference_angle, out, cap,
                               n_out, raw_to_plane, tier_ends);
    });
}

ixg_status ixg_asc_planes(const ixg_asc* s, ixg_facet_plane* out, uint32_t cap, uint32_t* n_out) {
    if (n_out) *n_out = 0;
    if (!s || !n_out) return IXG_INVALID;
    return guarded([&] { return schedule_planes_of(s, out, cap, n_out); });
}

}  // extern "C"
