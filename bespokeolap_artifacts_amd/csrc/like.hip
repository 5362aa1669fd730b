// like.hip — host-side ILIKE over a string dictionary (no device code; the
// device form over a dictionary in HBM is in ceb_q10a.hip).
//
// The reference evaluates ILIKE once per distinct string id and caches the
// answer (bespoke_ceb/query_q10a.cpp:43-209, 756-808); the dictionary-code
// columns then filter through a byte mask.  Here the mask over the whole
// dictionary is built on the host and the device reads it as a bitmap.
//
// Matching follows the reference's observable behaviour: both sides are
// lower-cased (ASCII); a pattern without '_' whose '%' signs sit only at the
// ends is a whole-string, prefix, suffix or substring test; any other
// pattern goes through the general scan, where '_' matches one character,
// '%' any run, and a pattern character equal to the text character
// (including a literal '%' facing a '%' in the text) is consumed as a match
// before '%' is considered.  On a mismatch the scan returns to the last '%'
// and retries one character further (the linear-space wildcard algorithm).
#include <cctype>
#include <cstring>
#include <string>

#include "common.hpp"

namespace {

char lower(char c) { return static_cast<char>(std::tolower(static_cast<unsigned char>(c))); }

bool wildcard(const std::string& p, const char* t, size_t tn) {
    size_t i = 0, j = 0;
    size_t star = std::string::npos, resume = 0;
    while (j < tn) {
        const char c = lower(t[j]);
        if (i < p.size() && (p[i] == '_' || p[i] == c)) {
            ++i;
            ++j;
        } else if (i < p.size() && p[i] == '%') {
            star = i++;
            resume = j;
        } else if (star != std::string::npos) {
            i = star + 1;
            j = ++resume;
        } else {
            return false;
        }
    }
    while (i < p.size() && p[i] == '%') ++i;
    return i == p.size();
}

using Kind = bolap::LikeKind;
using Pattern = bolap::LikePattern;

}  // namespace

bolap::LikePattern bolap::like_classify(const char* raw) {
    Pattern r;
    r.p = raw;
    for (char& c : r.p) c = lower(c);
    const std::string& p = r.p;
    if (p == "%") {
        r.kind = Kind::kAll;
        return r;
    }
    if (p.find('_') != std::string::npos) return r;
    const size_t first = p.find('%');
    if (first == std::string::npos) {
        r.kind = Kind::kExact;
        r.token = p;
        return r;
    }
    const size_t last = p.rfind('%');
    const size_t second = p.find('%', first + 1);
    if (first == 0 && last == p.size() - 1 && second == last) {
        r.kind = Kind::kContains;
        if (p.size() > 2) r.token = p.substr(1, p.size() - 2);
    } else if (first == 0 && second == std::string::npos) {
        r.kind = Kind::kSuffix;
        r.token = p.substr(1);
    } else if (last == p.size() - 1 && first == last) {
        r.kind = Kind::kPrefix;
        r.token = p.substr(0, p.size() - 1);
    }
    return r;
}

namespace {

bool matches(const Pattern& pt, const char* t, size_t tn) {
    std::string v(t, tn);
    for (char& c : v) c = lower(c);
    const std::string& k = pt.token;
    switch (pt.kind) {
        case Kind::kAll: return true;
        case Kind::kExact: return v == k;
        case Kind::kPrefix: return v.size() >= k.size() && v.compare(0, k.size(), k) == 0;
        case Kind::kSuffix:
            return v.size() >= k.size() && v.compare(v.size() - k.size(), k.size(), k) == 0;
        case Kind::kContains: return k.empty() || v.find(k) != std::string::npos;
        case Kind::kGeneral: return wildcard(pt.p, v.data(), v.size());
    }
    return false;
}

}  // namespace

extern "C" int bolap_ilike_mask(const char* chars, const int64_t* offsets, int64_t n,
                                const char* pattern, uint8_t* mask) {
    BOLAP_ARG_CHECK(n >= 0 && pattern && (n == 0 || (chars && offsets && mask)),
                    "bolap_ilike_mask: bad arguments");
    const Pattern pt = bolap::like_classify(pattern);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = offsets[i], b = offsets[i + 1];
        BOLAP_ARG_CHECK(0 <= a && a <= b, "bolap_ilike_mask: offsets not increasing at %lld",
                        (long long)i);
        mask[i] = matches(pt, chars + a, size_t(b - a)) ? 1 : 0;
    }
    return BOLAP_OK;
}
