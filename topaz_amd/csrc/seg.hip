// SEG low-complexity masking of query sequences, as TOPAZ applies it before
// a search (engine.c:254-263, run_seg in seg_wrapper.c:15-49): Wootton and
// Federhen's SEG with TOPAZ's settings (seg/seg.c:29-42: window 12, locut
// 2.2, hicut 2.5, maxtrim 100, no overlaps, hilenmin 0), each segment found
// overwritten with 'X'.  Host code: a few passes over one sequence.
//
// The rule, in the reference's terms:
//  * H[i] is the compositional entropy (bits) of the 12-letter window
//    starting at i - 5, counting only the 20 standard amino-acid letters
//    (either case); positions without a full window have none.
//  * Scanning i left to right, a trigger window H[i] <= 2.2 extends left
//    and right while H <= 2.5 (not past the last segment's end); the
//    extension is trimmed to its most improbable sub-window (of at most
//    100 fewer letters), the probability being that of its composition
//    among random sequences of 20 equiprobable letters; a trigger left of
//    the trimmed segment is searched again on the part it left out; the
//    scan resumes after the segment.
//  * Overlapping segments merge (seg/seg.c:410-453).
// Floating point follows the reference term for term: entropies and
// probabilities sum over the composition's nonzero counts in decreasing
// order, with ln(n!) at the 6 decimals the reference's table holds
// (seg/lnfac.h).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "../../include/topaz_amd.h"

namespace tpz {
namespace {

constexpr int kWin = 12;
constexpr int kDown = (kWin + 1) / 2 - 1;     // 5
constexpr int kUp = kWin - kDown;             // 7
constexpr double kLoCut = 2.2, kHiCut = 2.5;
constexpr int kMaxTrim = 100;
constexpr double kLn2 = 0.69314718055994530941723212145818;
constexpr double kLn20 = 2.9957322735539909;

// ln(n!) for n up to `n`, each rounded to 6 decimals.
void grow_lnfac(std::vector<double>& lnfac, size_t n) {
    char buf[64];
    for (size_t k = lnfac.size(); k < n; ++k) {
        snprintf(buf, sizeof buf, "%.6Lf", lgammal((long double)k + 1));
        lnfac.push_back(strtod(buf, nullptr));
    }
}

struct SegTables {
    int8_t cls[256];              // letter -> class 0..19, -1 for any other byte
    double entray[kWin + 1];      // per count in a full window: -x log2 x, x = count / 12
    std::vector<double> lnfac;    // ln(n!) rounded to 6 decimals, n <= 10,000
    SegTables() {
        static const char kLetters[] = "ACDEFGHIKLMNPQRSTVWY";
        for (int c = 0; c < 256; ++c) cls[c] = -1;
        for (int k = 0; k < 20; ++k) {
            cls[(unsigned char)kLetters[k]] = (int8_t)k;
            cls[(unsigned char)(kLetters[k] + 32)] = (int8_t)k;
        }
        entray[0] = 0;
        for (int i = 1; i <= kWin; ++i) {
            const double x = i / (double)kWin;
            entray[i] = -x * log(x) / kLn2;
        }
        grow_lnfac(lnfac, 10001);
    }
};

const SegTables& tables() {
    static const SegTables t;
    return t;
}

// The nonzero counts of a composition, decreasing (the reference's state
// vector); returns how many.
int state_of(const int32_t* comp, int32_t* sv) {
    int n = 0;
    for (int k = 0; k < 20; ++k)
        if (comp[k]) sv[n++] = comp[k];
    std::sort(sv, sv + n, [](int32_t a, int32_t b) { return a > b; });
    return n;
}

double entropy_of(const int32_t* comp, const SegTables& T) {
    int32_t sv[20];
    const int n = state_of(comp, sv);
    int total = 0;
    for (int k = 0; k < n; ++k) total += sv[k];
    double ent = 0.0;
    if (total == kWin) {
        for (int k = 0; k < n; ++k) ent += T.entray[sv[k]];
        return ent;
    }
    if (total == 0) return 0.0;
    const double r = 1.0 / (double)total;
    for (int k = 0; k < n; ++k) {
        const double x = sv[k];
        ent += x * log(x * r);
    }
    return -ent * r / kLn2;
}

// ln of the probability of a window of `total` letters with this
// composition: the assignments of the counts to classes, times the
// arrangements of the letters, over 20^total.
double log_prob(const int32_t* comp, int total, const std::vector<double>& lnfac) {
    int32_t sv[20];
    const int n = state_of(comp, sv);
    double ass = lnfac[20];
    if (n > 0) {
        for (int k = 0; k < n;) {
            int run = 1;
            while (k + run < n && sv[k + run] == sv[k]) ++run;
            ass -= lnfac[run];
            k += run;
        }
        if (n < 20) ass -= lnfac[20 - n];
    }
    double perm = lnfac[total];
    for (int k = 0; k < n; ++k) perm -= lnfac[sv[k]];
    return ass + perm - (double)total * kLn20;
}

void add(int32_t* comp, int8_t c, int d) {
    if (c >= 0) comp[c] += d;
}

// The most improbable sub-window of cls[0..n) at least n - 100 letters long
// (seg/seg.c:296-336): [lend, rend].
void trim(const int8_t* cls, int n, int& lend, int& rend, const std::vector<double>& lnfac) {
    lend = 0;
    rend = n - 1;
    int minlen = 1;
    if (n - kMaxTrim > minlen) minlen = n - kMaxTrim;
    double minprob = 1.0;
    for (int len = n; len > minlen; --len) {
        int32_t comp[20] = {0};
        for (int k = 0; k < len; ++k) add(comp, cls[k], 1);
        for (int i = 0;; ++i) {
            const double p = log_prob(comp, len, lnfac);
            if (p < minprob) {
                minprob = p;
                lend = i;
                rend = len + i - 1;
            }
            if (i + len >= n) break;
            add(comp, cls[i], -1);
            add(comp, cls[i + len], 1);
        }
    }
}

struct Seg {
    int begin, end;
};

// seg/seg.c:152-205 over cls[0..n), positions reported + offset.
void segment(const int8_t* cls, int n, int offset, std::vector<Seg>& out, const SegTables& T,
             const std::vector<double>& lnfac) {
    if (kWin > n) return;
    std::vector<double> H((size_t)n, -1.0);
    {
        int32_t comp[20] = {0};
        for (int k = 0; k < kWin; ++k) add(comp, cls[k], 1);
        for (int i = kDown; i <= n - kUp; ++i) {
            H[i] = entropy_of(comp, T);
            const int s = i - kDown;            // the window [s, s + 12) moves on
            if (s + kWin < n) {
                add(comp, cls[s], -1);
                add(comp, cls[s + kWin], 1);
            }
        }
    }
    const int first = kDown, last = n - kUp;
    int lowlim = first;
    for (int i = first; i <= last; ++i) {
        if (!(H[i] <= kLoCut && H[i] != -1)) continue;
        int lo = i;
        while (lo >= lowlim && H[lo] != -1 && !(H[lo] > kHiCut)) --lo;
        const int loi = lo + 1;
        int hi = i;
        while (hi <= last && H[hi] != -1 && !(H[hi] > kHiCut)) ++hi;
        const int hii = hi - 1;
        int leftend = loi - kDown, rightend = hii + kUp - 1;
        int lend, rend;
        trim(cls + leftend, rightend - leftend + 1, lend, rend, lnfac);
        const int len0 = rightend - leftend + 1;
        leftend += lend;
        rightend -= len0 - rend - 1;
        if (i + kUp - 1 < leftend) {
            // a trigger window in what the trim cut off on the left
            const int l0 = loi - kDown, r0 = leftend - 1;
            segment(cls + l0, r0 - l0 + 1, offset + l0, out, T, lnfac);
        }
        out.push_back(Seg{leftend + offset, rightend + offset});
        i = std::min(hii, rightend + kDown);
        lowlim = i + 1;
    }
}

}  // namespace
}  // namespace tpz

extern "C" int32_t tpz_seg(char* seq, int32_t len) {
    using namespace tpz;
    if ((!seq && len > 0) || len < 0) return -1;
    const SegTables& T = tables();
    // ln(n!) up to the longest window a trim can take: beyond the shared
    // table (n <= 10,000) a copy of this call's own
    std::vector<double> longer;
    if (T.lnfac.size() < (size_t)len + 21) {
        longer = T.lnfac;
        grow_lnfac(longer, (size_t)len + 21);
    }
    const std::vector<double>& lnfac = longer.empty() ? T.lnfac : longer;
    std::vector<int8_t> cls((size_t)len);
    for (int32_t k = 0; k < len; ++k) cls[k] = T.cls[(unsigned char)seq[k]];
    std::vector<Seg> segs;
    segment(cls.data(), len, 0, segs, T, lnfac);
    // merge overlapping segments (later ones take over the end)
    std::vector<Seg> merged;
    for (const Seg& s : segs) {
        if (!merged.empty() && merged.back().end >= s.begin) merged.back().end = s.end;
        else merged.push_back(s);
    }
    std::vector<uint8_t> hit((size_t)len, 0);
    for (const Seg& s : merged)
        for (int k = s.begin; k <= s.end; ++k) hit[k] = 1;
    int32_t masked = 0;
    for (int32_t k = 0; k < len; ++k)
        if (hit[k]) {
            seq[k] = 'X';
            ++masked;
        }
    return masked;
}
