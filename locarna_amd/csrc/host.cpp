// Host preprocessing for the MI355X aligner (see host.hpp).
#include "host.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <sstream>
#include <thread>

#include "ribosum85_60.h"

namespace locarna_amd {

// ------------------------------------------------------------------ RnaData

namespace {
// A line that is neither empty nor starts with white space (white-space
// lines are comments in pp 2.0), trailing blanks removed.
bool next_content_line(std::istream &in, std::string &line) {
    while (std::getline(in, line)) {
        if (!line.empty() && !std::isspace((unsigned char)line[0])) {
            size_t e = line.find_last_not_of(" \t\r");
            line.resize(e + 1);
            return true;
        }
    }
    return false;
}
bool starts_with(const std::string &s, const char *p) { return s.compare(0, std::strlen(p), p) == 0; }
}  // namespace

RnaData::RnaData(const std::string &pp_text, double p_bpcut, long max_bp_span, bool stacking) : p_bpcut_(p_bpcut) {
    std::istringstream in(pp_text);
    std::string line;
    std::getline(in, line);
    if (!starts_with(line, "#PP 2")) throw failure("RnaData: input is not in #PP 2.0 format");
    std::vector<std::string> names;
    std::map<std::string, std::string> rows;
    while (next_content_line(in, line)) {
        if (starts_with(line, "#END")) break;
        if (line[0] == '#') continue;  // #A anchors, #S structure: not used by this path
        std::istringstream ls(line);
        std::string nm, s;
        ls >> nm >> s;
        if (!rows.count(nm)) names.push_back(nm);
        rows[nm] += s;
    }
    if (names.empty()) throw failure("RnaData: no sequence in the pp alignment block");
    for (const std::string &nm : names) {
        std::string r = rows[nm];
        for (auto &c : r) {  // aux.cc:65-70
            c = (char)std::toupper((unsigned char)c);
            if (c == 'T') c = 'U';
        }
        if (!rows_.empty() && r.size() != rows_[0].size())
            throw failure("RnaData: rows of the pp alignment differ in length");
        rows_.push_back(r);
    }
    names_ = names;
    name_ = names[0];
    seq_ = rows_[0];
    if (!next_content_line(in, line) || line != "#SECTION BASEPAIRS")
        throw failure("Syntax error: Expected base pair section header.");
    const size_t span_max = max_bp_span < 0 ? std::numeric_limits<size_t>::max() : (size_t)max_bp_span;
    bool stack_keyword = false;
    while (next_content_line(in, line)) {
        if (line[0] == '#') {
            if (starts_with(line, "#END")) break;
            if (starts_with(line, "#STACK")) stack_keyword = true;
            if (starts_with(line, "#BPCUT")) {
                std::istringstream ls(line);
                std::string kw;
                double p;
                ls >> kw >> p;
                if (ls.fail()) throw failure("Syntax error: Cannot parse line \"" + line + "\" in base pairs section.");
                p_bpcut_ = std::max(p, p_bpcut_);
            }
            continue;
        }
        std::istringstream ls(line);
        size_t i, j;
        double p;
        ls >> i >> j >> p;
        if (ls.fail()) throw failure("Syntax error: Cannot parse line \"" + line + "\" in base pairs section.");
        if (!(1 <= i && i < j && j <= seq_.size()))
            throw failure("Syntax error: Invalid indices in PP input line \"" + line + "\".");
        if (p <= p_bpcut_ || j - i + 1 > span_max) continue;
        probs_[{(int)i, (int)j}] = p;
        // with stacking a fourth column: the joint probability of (i, j) and
        // (i+1, j-1), kept above the cut (rna_data.cc:1083-1092)
        double p2;
        if (stacking && (ls >> p2) && p2 > p_bpcut_) joint_[{(int)i, (int)j}] = p2;
    }
    if (!stack_keyword && !joint_.empty())  // rna_data.cc:1097-1100
        throw failure("Syntax error: Stacking probabilties found but stack keyword is missing.");
    for (const auto &x : probs_) pairs_.push_back(x);
    // paired probabilities per position, summed in the reference's orders
    // (upstream: j ascending; downstream: i ascending): both are the order
    // of the map, (i, j) ascending
    up_.assign(seq_.size() + 2, 0.0);
    down_.assign(seq_.size() + 2, 0.0);
    for (const auto &x : probs_) {
        up_[x.first.first] += x.second;
        down_[x.first.second] += x.second;
    }
}

double RnaData::joint_arc_prob(int i, int j) const {
    auto it = joint_.find({i, j});
    return it == joint_.end() ? 0.0 : it->second;
}

double RnaData::arc_prob(int i, int j) const {
    auto it = probs_.find({i, j});
    return it == probs_.end() ? 0.0 : it->second;
}

double RnaData::prob_paired_upstream(int i) const { return up_[i]; }

double RnaData::prob_paired_downstream(int j) const { return down_[j]; }

// ---------------------------------------------------------------- BasePairs

BasePairs::BasePairs(const RnaData *rna, double min_prob) : len_(rna->length()) {
    left_.resize(len_ + 2);
    right_.resize(len_ + 2);
    // arc index order: left end from len-3 down to 1, right end ascending
    // (basepairs.cc:162-177)
    std::vector<std::vector<std::pair<int, double>>> by_left(len_ + 1);
    for (const auto &x : rna->pairs()) {
        int i = x.first.first, j = x.first.second;
        if (j >= i + 3) by_left[i].push_back({j, x.second});
    }
    for (int i = len_ - 3; i >= 1; --i) {
        for (const auto &jp : by_left[i]) {  // ascending j
            if (jp.second >= min_prob) {
                int idx = (int)arcs_.size();
                arcs_.push_back(Arc{idx, i, jp.first});
                left_[i].push_back(idx);
                right_[jp.first].push_back(idx);
            }
        }
    }
    for (auto &l : left_)
        std::sort(l.begin(), l.end(), [&](int a, int b) { return arcs_[a].right < arcs_[b].right; });
    for (auto &l : right_)
        std::sort(l.begin(), l.end(), [&](int a, int b) { return arcs_[a].left > arcs_[b].left; });
}

BasePairs::BasePairs(int len, const std::vector<std::pair<int, int>> &sorted_arcs) : len_(len) {
    left_.resize(len_ + 2);
    right_.resize(len_ + 2);
    for (const auto &x : sorted_arcs) {
        const int idx = (int)arcs_.size();
        arcs_.push_back(Arc{idx, x.first, x.second});
        left_[x.first].push_back(idx);
        right_[x.second].push_back(idx);
    }
    for (auto &l : left_)
        std::sort(l.begin(), l.end(), [&](int a, int b) { return arcs_[a].right < arcs_[b].right; });
    for (auto &l : right_)
        std::sort(l.begin(), l.end(), [&](int a, int b) { return arcs_[a].left > arcs_[b].left; });
}

int BasePairs::find(int i, int j) const {
    if (i < 0 || i > len_) return -1;
    for (int a : left_[i])
        if (arcs_[a].right == j) return a;
    return -1;
}

std::vector<ExplicitArcMatch> parse_arcmatch_scores(const std::string &text, int lenA, int lenB,
                                                    int probability_scale) {
    std::vector<ExplicitArcMatch> out;
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        long i = 0, j = 0, k = 0, l = 0;
        ExplicitArcMatch x;
        bool ok = static_cast<bool>(ls >> i >> j >> k >> l);
        if (ok) {
            if (probability_scale < 0) {
                long sc = 0;
                ok = static_cast<bool>(ls >> sc);
                x.score = sc;
            } else {
                double prob = 0;
                ok = static_cast<bool>(ls >> prob);
                x.score = (score_t)(prob * (double)probability_scale);
            }
        }
        if (!ok || i <= 0 || j <= 0 || k <= 0 || l <= 0 || i > j || j > lenA || k > l || l > lenB)
            throw failure("Cannot read arc match scores. Invalid line " + std::to_string(lineno) + ": " + line);
        if (j - i < 3 || l - k < 3)
            throw failure("Cannot read arc match scores. Arc shorter than a base pair of BasePairs (j - i < 3) "
                          "in line " + std::to_string(lineno) + ": " + line);
        x.i = (int)i;
        x.j = (int)j;
        x.k = (int)k;
        x.l = (int)l;
        out.push_back(x);
    }
    return out;
}

// ---------------------------------------------------------- TraceController

TraceController::TraceController(int lenA, int lenB)
    : lenA_(lenA), lenB_(lenB), min_col_(lenA + 1, 0), max_col_(lenA + 1, lenB) {}

void TraceController::constrain_wo_ref(int delta) {
    // unsigned arithmetic as the reference (size_type)
    const size_t la = lenA_, lb = lenB_, d = (size_t)delta;
    if (la == 0 || lb == 0) return;
    for (size_t i = 0; i <= la; i++) {
        const size_t x = i * lb * (la + lb);
        size_t y = 2 * d * la * lb;
        const size_t z = la * (la + lb);
        if (la > lb)
            y = std::max(y, (la + lb) * la / 2);
        else if (lb > la)
            y = std::max(y, (la + lb) * lb / 2);
        min_col_[i] = (int)(x > y ? (x - y + z - 1) / z : 0);
        max_col_[i] = (int)std::min((x + y) / z, lb);
    }
}

void TraceController::restrict_by_trace_probabilities(const std::vector<double> &probs, double min_prob) {
    const int W = lenB_ + 1;
    std::vector<int> first(lenA_ + 1, -1), last(lenA_ + 1, -1);
    for (int i = 0; i <= lenA_; ++i) {
        for (int j = min_col_[i]; j <= max_col_[i]; ++j) {
            if (probs[(size_t)i * W + j] >= min_prob) {
                if (first[i] < 0) first[i] = j;
                last[i] = j;
            }
        }
    }
    restrict_by_rows(first, last);
}

void TraceController::restrict_by_rows(const std::vector<int> &first, const std::vector<int> &last) {
    for (int i = 0; i <= lenA_; ++i) {
        // no passing column: the row's range becomes empty (max < min)
        const int lo = first[i] >= 0 ? std::min(first[i], max_col_[i]) : max_col_[i];
        const int hi = last[i] >= 0 ? std::max(last[i], min_col_[i]) : min_col_[i];
        min_col_[i] = std::max(min_col_[i], lo);
        max_col_[i] = std::min(max_col_[i], hi);
    }
    int run = 0;
    for (int i = 0; i <= lenA_; ++i) run = max_col_[i] = std::max(max_col_[i], run);
    run = max_col_[lenA_];
    for (int i = lenA_; i >= 0; --i) run = min_col_[i] = std::min(min_col_[i], run);
}

// ----------------------------------------------------------- band (PFGotoh)

int nt_code(char c) {
    switch (c) {
        case 'A': return 0;
        case 'C': return 1;
        case 'G': return 2;
        case 'U': return 3;
    }
    return -1;
}

void band_similarity(const ScoringParams &sp, double sim[16]) {
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            sim[a * 4 + b] = sp.use_ribosum ? ribosum85_60::bm[a * 4 + b] : (a == b ? sp.match : sp.mismatch);
}

namespace {

// Gotoh partition function over cells (i, j) of the band, in the given
// direction.  Forward: recursion from (0,0) with similarity sim(i, j).
// Backward: the same recursion on the reversed sequences; written here on
// original coordinates, Xr(i, j) holds the reference's X_rev(lenA-i,
// lenB-j), whose similarity at reversed cell (lenA-i, lenB-j) equals
// sim(i+1, j+1) bit for bit (sqrt terms added in swapped order, which is
// exact).  Operation order follows edge_probs.icc:76-162.
template <class T>
struct GotohPF {
    int n, m;
    bool local;
    double g_open, g_ext;
    int free_endgaps;  // LCG_FREE_* bits of this orientation
    std::vector<T> M, A, B;  // (n+1)*(m+1)
    T &at(std::vector<T> &v, int i, int j) { return v[(size_t)i * (m + 1) + j]; }
};

template <class T>
void gotoh_forward(GotohPF<T> &g, const std::vector<double> &match_w, const std::vector<int> &mn,
                   const std::vector<int> &mx) {
    const int n = g.n, m = g.m;
    const double go = g.g_open, ge = g.g_ext;
    g.M.assign((size_t)(n + 1) * (m + 1), (T)0);
    g.A = g.M;
    g.B = g.M;
    auto valid = [&](int i, int j) { return mn[i] <= j && j <= mx[i]; };
    if (valid(0, 0)) g.at(g.M, 0, 0) = g.local ? 0 : 1;
    if (n > 0 && valid(1, 0)) g.at(g.A, 1, 0) = go * ge;
    if (m > 0 && valid(0, 1)) g.at(g.B, 0, 1) = go * ge;
    for (int i = 2; i <= n; ++i) {
        if (mn[i] > 0) break;
        g.at(g.A, i, 0) = g.at(g.A, i - 1, 0) * ge;
    }
    for (int j = std::max(mn[0], 2); j <= std::min(mx[0], m); ++j) g.at(g.B, 0, j) = g.at(g.B, 0, j - 1) * ge;
    // free end gaps (edge_probs.icc:120-131): left end of the second
    // sequence (bit 4) on column 0, of the first (bit 1) on row 0
    if (g.free_endgaps & 4)
        for (int i = 1; i <= n; ++i) g.at(g.A, i, 0) += 1;
    if (g.free_endgaps & 1)
        for (int j = 1; j <= m; ++j) g.at(g.B, 0, j) += 1;
    const double lw = 0;
    for (int i = 1; i <= n; ++i) {
        for (int j = std::max(mn[i], 1); j <= std::min(mx[i], m); ++j) {
            double w = match_w[(size_t)i * (m + 1) + j];
            g.at(g.M, i, j) = g.at(g.M, i - 1, j - 1) * w + g.at(g.A, i - 1, j - 1) * w +
                              g.at(g.B, i - 1, j - 1) * w + (g.local ? w : lw);
            g.at(g.A, i, j) = g.at(g.A, i - 1, j) * ge + g.at(g.M, i - 1, j) * go * ge + g.at(g.B, i - 1, j) * go * ge;
            g.at(g.B, i, j) = g.at(g.B, i, j - 1) * ge + g.at(g.M, i, j - 1) * go * ge + g.at(g.A, i, j - 1) * go * ge;
        }
    }
}

template <class T>
std::vector<double> trace_probs(const RnaData &RA, const RnaData &RB, const TraceController &tc,
                                const ScoringParams &sp, bool local, int fe) {
    const int n = RA.length(), m = RB.length();
    const double temp = sp.temperature_alipf / 100.0;
    const double gap_open = sp.indel_opening / 100.0, gap_ext = sp.indel / 100.0;
    const double sw = sp.struct_weight / 100.0;
    double simv[16];
    band_similarity(sp, simv);
    auto sim = [&](int a, int b) { return simv[a * 4 + b]; };
    std::vector<double> upA(n + 1, 0.0), dnA(n + 1, 0.0), upB(m + 1, 0.0), dnB(m + 1, 0.0);
    for (int i = 1; i <= n; ++i) {
        upA[i] = RA.prob_paired_upstream(i);
        dnA[i] = RA.prob_paired_downstream(i);
    }
    for (int j = 1; j <= m; ++j) {
        upB[j] = RB.prob_paired_upstream(j);
        dnB[j] = RB.prob_paired_downstream(j);
    }
    // Boltzmann weights of the similarity, forward orientation; the
    // sequence term averages over the row pairs of the two columns whose
    // symbols are both in ACGU, summed in row order (stral_score.cc:29-60)
    const int rA = RA.num_rows(), rB = RB.num_rows();
    std::vector<int> cA((size_t)rA * (n + 1)), cB((size_t)rB * (m + 1));
    for (int k = 0; k < rA; ++k)
        for (int i = 1; i <= n; ++i) cA[(size_t)k * (n + 1) + i] = nt_code(RA.row(k)[i - 1]);
    for (int l = 0; l < rB; ++l)
        for (int j = 1; j <= m; ++j) cB[(size_t)l * (m + 1) + j] = nt_code(RB.row(l)[j - 1]);
    std::vector<double> fw((size_t)(n + 1) * (m + 1), 0.0);
    // rows of the similarity on several threads for large profile pairs
    // (each cell's sum keeps its row order, so the values do not change)
    const double work = (double)rA * rB * n * m;
    const int nth = (rA * rB > 1 && work >= 4e6)
                        ? (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()))
                        : 1;
    auto rows = [&](int t) {
    for (int i = 1 + t; i <= n; i += nth) {
        for (int j = 1; j <= m; ++j) {
            double seq = 0;
            int cnt = 0;
            for (int k = 0; k < rA; ++k) {
                const int a = cA[(size_t)k * (n + 1) + i];
                for (int l = 0; l < rB; ++l) {
                    const int b = cB[(size_t)l * (m + 1) + j];
                    if (a >= 0 && b >= 0) {
                        seq += sim(a, b);
                        cnt++;
                    }
                }
            }
            if (cnt != 0) seq /= cnt;
            double s = sw * (std::sqrt(dnA[i] * dnB[j]) + std::sqrt(upA[i] * upB[j])) + seq;
            fw[(size_t)i * (m + 1) + j] = std::exp(s / temp);
        }
    }
    };
    if (nth == 1) {
        rows(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t) pool.emplace_back(rows, t);
        for (auto &th : pool) th.join();
    }
    GotohPF<T> F{n, m, local, std::exp(gap_open / temp), std::exp(gap_ext / temp), fe};
    gotoh_forward(F, fw, tc.min_cols(), tc.max_cols());
    // reversed problem: cell (i', j') of the reversed matrices <-> (n-i', m-j')
    std::vector<double> rw((size_t)(n + 1) * (m + 1), 0.0);
    for (int ip = 1; ip <= n; ++ip)
        for (int jp = 1; jp <= m; ++jp) rw[(size_t)ip * (m + 1) + jp] = fw[(size_t)(n + 1 - ip) * (m + 1) + (m + 1 - jp)];
    std::vector<int> rmn(n + 1), rmx(n + 1);
    for (int i = 0; i <= n; ++i) {
        rmx[n - i] = m - tc.min_col(i);
        rmn[n - i] = m - tc.max_col(i);
    }
    // the reversed sequences' ends: left and right swapped (FreeEndgaps::reverse)
    const int rfe = ((fe & 2) >> 1) | ((fe & 1) << 1) | ((fe & 8) >> 1) | ((fe & 4) << 1);
    GotohPF<T> R{n, m, local, F.g_open, F.g_ext, rfe};
    gotoh_forward(R, rw, rmn, rmx);
    T z;
    if (local) {
        z = 1;
        for (int i = 0; i <= n; ++i)
            for (int j = 0; j <= m; ++j) z += F.at(F.M, i, j);
    } else {
        z = F.at(F.M, n, m) + F.at(F.A, n, m) + F.at(F.B, n, m);
        // edge_probs.icc:39-53: under the forward left flags
        if (fe & 4)
            for (int i = 0; i <= n; ++i) z += F.at(F.A, i, m);
        if (fe & 1)
            for (int j = 0; j <= m; ++j) z += F.at(F.B, n, j);
    }
    std::vector<double> probs((size_t)(n + 1) * (m + 1), 0.0);
    const double locality_add = local ? 1 : 0;
    const double g_open = std::exp(gap_open / temp);
    for (int i = 0; i <= n; ++i) {
        for (int j = std::max(tc.min_col(i), 0); j <= std::min(tc.max_col(i), m); ++j) {
            int ri = n - i, rj = m - j;
            T rM = R.at(R.M, ri, rj), rA = R.at(R.A, ri, rj), rB = R.at(R.B, ri, rj);
            T zij = F.at(F.M, i, j) * (rM + rA + rB + locality_add) + F.at(F.A, i, j) * (rM + rA / g_open + rB) +
                    F.at(F.B, i, j) * (rM + rA + rB / g_open);
            probs[(size_t)i * (m + 1) + j] = zij / z;
        }
    }
    return probs;
}

}  // namespace

void restrict_trace_by_probabilities(TraceController &tc, const RnaData &A, const RnaData &B,
                                     const ScoringParams &sp, double min_trace_prob, bool extended_precision,
                                     bool local, int free_endgaps) {
    if (!(min_trace_prob > 0.0)) return;
    std::vector<double> probs = extended_precision ? trace_probs<long double>(A, B, tc, sp, local, free_endgaps)
                                                   : trace_probs<double>(A, B, tc, sp, local, free_endgaps);
    tc.restrict_by_trace_probabilities(probs, min_trace_prob);
}

// --------------------------------------------------------------- ArcMatches

ArcMatches::ArcMatches(const BasePairs &A, const BasePairs &B, const TraceController &tc, int max_diff_am,
                       int max_diff_at_am)
    : A_(A), B_(B) {
    const int lmax = std::max(A.seqlen(), B.seqlen());
    const int dlen = max_diff_am < 0 ? lmax : max_diff_am;
    const int dend = max_diff_at_am < 0 ? lmax : max_diff_at_am;
    for (int a = 0; a < A.num_bps(); ++a) {
        const Arc &x = A.arc(a);
        for (int b = 0; b < B.num_bps(); ++b) {
            const Arc &y = B.arc(b);
            bool ok = tc.is_valid_match(x.left, y.left) && tc.is_valid_match(x.right, y.right) &&
                      std::abs((x.right - x.left) - (y.right - y.left)) <= dlen &&
                      std::abs(x.left - y.left) <= dend && std::abs(x.right - y.right) <= dend;
            if (ok) ams_.push_back(ArcMatch{(int)ams_.size(), a, b});
        }
    }
}

ArcMatches::ArcMatches(const BasePairs &A, const BasePairs &B, const TraceController &tc,
                       const std::vector<ExplicitArcMatch> &lines, int max_diff_am, int max_diff_at_am)
    : A_(A), B_(B), explicit_(true) {
    const int lmax = std::max(A.seqlen(), B.seqlen());
    const int dlen = max_diff_am < 0 ? lmax : max_diff_am;
    const int dend = max_diff_at_am < 0 ? lmax : max_diff_at_am;
    for (const ExplicitArcMatch &x : lines) {
        const int a = A.find(x.i, x.j), b = B.find(x.k, x.l);
        const Arc &p = A.arc(a), &q = B.arc(b);
        // is_valid_arcmatch (arc_matches.cc:19-48)
        const bool ok = tc.is_valid_match(p.left, q.left) && tc.is_valid_match(p.right, q.right) &&
                        std::abs((p.right - p.left) - (q.right - q.left)) <= dlen &&
                        std::abs(p.left - q.left) <= dend && std::abs(p.right - q.right) <= dend;
        if (!ok) continue;
        ams_.push_back(ArcMatch{(int)ams_.size(), a, b});
        scores_.push_back(x.score);
    }
}

// ------------------------------------------------------------------ Scoring

Scoring::Scoring(const RnaData &A, const RnaData &B, const ArcMatches &am, const ScoringParams &p)
    : p_(p), am_(am) {
    // symbol codes: ACGU first, other symbols in order of appearance
    symbols_ = {'A', 'C', 'G', 'U'};
    auto code_of = [&](char c) {
        for (size_t k = 0; k < symbols_.size(); ++k)
            if (symbols_[k] == c) return (int)k;
        if ((int)symbols_.size() >= NCODES) throw failure("Scoring: too many distinct sequence symbols");
        symbols_.push_back(c);
        return (int)symbols_.size() - 1;
    };
    const int n = A.length(), m = B.length();
    lenB_ = m;
    profile_ = A.num_rows() > 1 || B.num_rows() > 1;
    // score of one pair of row symbols (scoring.cc:162-190): corrected
    // RIBOSUM log-odds for two nucleotides, otherwise match / mismatch
    // unless one of them is N (gap symbols compare as symbols)
    auto symbol_score = [&](char ca, char cb) -> score_t {
        const int xa = nt_code(ca), xb = nt_code(cb);
        if (p.use_ribosum && xa >= 0 && xb >= 0) {
            double bg = ribosum85_60::base_nonstruct[xa] * ribosum85_60::base_nonstruct[xb];
            double corrected = std::log(ribosum85_60::basematch[xa * 4 + xb] / bg) / std::log(2);
            return round2score(100.0 * corrected);
        }
        if (ca != 'N' && cb != 'N') return (ca == cb) ? p.match : p.mismatch;
        return 0;
    };
    codeA_.assign(n + 1, 0);
    codeB_.assign(m + 1, 0);
    sigma_codes_.assign(NCODES * NCODES, 0);
    if (profile_) {
        // sum of pairs over the rows of the two columns, integer quotient by
        // the number of row pairs (scoring.cc:140-198)
        const int rA = A.num_rows(), rB = B.num_rows();
        sigma_tab_.assign((size_t)(n + 1) * (m + 1), 0);
        std::vector<score_t> tab(256 * 256);
        for (int x = 0; x < 256; ++x)
            for (int y = 0; y < 256; ++y) tab[x * 256 + y] = symbol_score((char)x, (char)y);
        // the integer sum over row pairs equals the sum over pairs of
        // distinct symbols weighted by their counts in the two columns
        auto histogram = [](const RnaData &R, int len) {
            std::vector<std::vector<std::pair<unsigned char, score_t>>> h(len + 1);
            for (int i = 1; i <= len; ++i)
                for (int k = 0; k < R.num_rows(); ++k) {
                    const unsigned char c = (unsigned char)R.row(k)[i - 1];
                    bool found = false;
                    for (auto &e : h[i])
                        if (e.first == c) ++e.second, found = true;
                    if (!found) h[i].push_back({c, 1});
                }
            return h;
        };
        const auto hA = histogram(A, n), hB = histogram(B, m);
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= m; ++j) {
                score_t sum = 0;
                for (const auto &x : hA[i])
                    for (const auto &y : hB[j]) sum += x.second * y.second * tab[x.first * 256 + y.first];
                sigma_tab_[(size_t)i * (m + 1) + j] = round2score((double)(sum / (int)(rA * rB)));
            }
    } else {
        for (int i = 1; i <= n; ++i) codeA_[i] = code_of(A.nt(i));
        for (int j = 1; j <= m; ++j) codeB_[j] = code_of(B.nt(j));
        // sigma per symbol pair (scoring.cc:140-198, single-row alignments)
        for (size_t a = 0; a < symbols_.size(); ++a)
            for (size_t b = 0; b < symbols_.size(); ++b)
                sigma_codes_[a * NCODES + b] = round2score((double)(symbol_score(symbols_[a], symbols_[b]) / (int)1));
    }
    // position-specific gap cost from the column's gap frequency, in single
    // precision (scoring.cc:271-311)
    auto gapcost = [&](const RnaData &R, std::vector<score_t> &g) {
        g.assign(R.length() + 1, 0);
        for (int i = 1; i <= R.length(); ++i) {
            float freq = 0;
            for (int k = 0; k < R.num_rows(); ++k) freq += (R.row(k)[i - 1] == '-') ? 1 : 0;
            freq /= (size_t)R.num_rows();
            g[i] = round2score((1 - freq) * (score_t)p.indel);
        }
    };
    gapcost(A, gapA_);
    gapcost(B, gapB_);
    // arc weights (scoring.cc:200-265)
    const double peA = p.exp_prob > 0 ? p.exp_prob : 1.0 / (2.0 * n);
    const double peB = p.exp_prob > 0 ? p.exp_prob : 1.0 / (2.0 * m);
    auto weight = [&](double prob, double pe) {
        return round2score(std::round((score_t)p.struct_weight * (1 - std::log(prob) / std::log(pe))));
    };
    const BasePairs &bA = am.bpsA(), &bB = am.bpsB();
    wA_.resize(bA.num_bps());
    wB_.resize(bB.num_bps());
    const int na = am.num_arc_matches();
    am_score_.resize(na);
    am_seq_.assign(na, 0);
    if (am.explicit_scores()) {
        // explicit arc-match scores (Scoring::arcmatch, scoring.cc:541-546);
        // the arc weights take no part
        wA_.assign(bA.num_bps(), 0);
        wB_.assign(bB.num_bps(), 0);
        for (int k = 0; k < na; ++k) am_score_[k] = am.explicit_score(k);
        return;
    }
    for (int k = 0; k < bA.num_bps(); ++k) wA_[k] = weight(A.arc_prob(bA.arc(k).left, bA.arc(k).right), peA);
    for (int k = 0; k < bB.num_bps(); ++k) wB_[k] = weight(B.arc_prob(bB.arc(k).left, bB.arc(k).right), peB);
    // arc-match scores (scoring.cc:368-554)
    for (int k = 0; k < na; ++k) {
        const ArcMatch &x = am.arcmatch(k);
        const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
        score_t seq = 0;
        if (p.tau != 0) {
            if (p.use_ribosum && profile_) {
                seq = ribosum_arc_term_rows(A, a, B, b);
            } else if (p.use_ribosum) {
                int al = nt_code(A.nt(a.left)), ar = nt_code(A.nt(a.right));
                int bl = nt_code(B.nt(b.left)), br = nt_code(B.nt(b.right));
                if (al >= 0 && ar >= 0 && bl >= 0 && br >= 0) seq = ribosum_arc_term(al, ar, bl, br);
            } else {
                seq = basematch(a.left, b.left) + basematch(a.right, b.right);
            }
        }
        am_seq_[k] = seq;
        am_score_[k] = ((score_t)p.tau * seq) / 100 + (wA_[x.arcA] + wB_[x.arcB]);
    }
    if (!p.stacking) return;
    // stacked weights (scoring.cc:200-236): --stacking the weight of the
    // stacked probability p2 / p(i+1, j-1), --new-stacking the arc's weight
    // plus that of p2, both: the sum of the two terms; for arcs whose inner
    // arc exists and that have a joint probability (only arc matches of such
    // arcs on both sides use them; elsewhere the reference's weight of a zero
    // probability is not finite)
    auto stack_w = [&](const RnaData &R, const BasePairs &bp, const std::vector<score_t> &w, double pe) {
        std::vector<score_t> sw(bp.num_bps(), 0);
        for (int k = 0; k < bp.num_bps(); ++k) {
            const Arc &a = bp.arc(k);
            const double pin = R.arc_prob(a.left + 1, a.right - 1), p2 = R.joint_arc_prob(a.left, a.right);
            const bool ok = pin > 0 && p2 > 0;
            if ((p.stacking & 1) && ok) sw[k] = weight(p2 / pin, pe);
            if (p.stacking & 2) {
                if (!(p.stacking & 1)) sw[k] = w[k];
                if (ok) sw[k] += weight(p2, pe);
            }
        }
        return sw;
    };
    const std::vector<score_t> swA = stack_w(A, bA, wA_, peA), swB = stack_w(B, bB, wB_, peB);
    am_score_st_.assign(na, 0);
    am_stackable_.assign(na, 0);
    for (int k = 0; k < na; ++k) {
        const ArcMatch &x = am.arcmatch(k);
        const Arc &a = bA.arc(x.arcA), &b = bB.arc(x.arcB);
        am_stackable_[k] = A.joint_arc_prob(a.left, a.right) > 0 && B.joint_arc_prob(b.left, b.right) > 0;
        if (am_stackable_[k]) am_score_st_[k] = ((score_t)p.tau * am_seq_[k]) / 100 + (swA[x.arcA] + swB[x.arcB]);
    }
}

score_t Scoring::ribosum_arc_term(int al, int ar, int bl, int br) {
    double pam = ribosum85_60::arcmatch[(al * 4 + ar) * 16 + (bl * 4 + br)];
    double pa = ribosum85_60::basepair[al * 4 + ar], pb = ribosum85_60::basepair[bl * 4 + br];
    double sc = 0;
    sc += std::log(pam / (pa * pb)) / std::log(2);
    return round2score(100.0 * sc / 1);
}

score_t Scoring::ribosum_arc_term_rows(const RnaData &A, const Arc &a, const RnaData &B, const Arc &b) {
    double sc = 0;
    int considered = 0;
    for (int k = 0; k < A.num_rows(); ++k) {
        const int al = nt_code(A.row(k)[a.left - 1]), ar = nt_code(A.row(k)[a.right - 1]);
        for (int l = 0; l < B.num_rows(); ++l) {
            const int bl = nt_code(B.row(l)[b.left - 1]), br = nt_code(B.row(l)[b.right - 1]);
            if (al < 0 || ar < 0 || bl < 0 || br < 0) continue;
            ++considered;
            const double pam = ribosum85_60::arcmatch[(al * 4 + ar) * 16 + (bl * 4 + br)];
            const double pa = ribosum85_60::basepair[al * 4 + ar], pb = ribosum85_60::basepair[bl * 4 + br];
            sc += std::log(pam / (pa * pb)) / std::log(2);
        }
    }
    if (considered == 0) return 0;
    return round2score(100.0 * sc / considered);
}

score_t Scoring::max_abs_sigma() const {
    score_t mx = 0;
    for (score_t v : sigma_codes_) mx = std::max(mx, v < 0 ? -v : v);
    for (score_t v : sigma_tab_) mx = std::max(mx, v < 0 ? -v : v);
    return mx;
}

score_t Scoring::basematch(int i, int j) const {
    if (profile_) return sigma_tab_[(size_t)i * (lenB_ + 1) + j];
    return sigma_codes_[codeA_[i] * NCODES + codeB_[j]];
}

// ------------------------------------------------------------ writers

std::vector<std::pair<std::string, std::string>> aligned_rows(const RnaData &A, const RnaData &B,
                                                              const std::vector<std::pair<int, int>> &cols) {
    bool clash = false;
    for (int k = 0; k < A.num_rows() && !clash; ++k)
        for (int l = 0; l < B.num_rows() && !clash; ++l) clash = A.row_name(k) == B.row_name(l);
    std::vector<std::pair<std::string, std::string>> out;
    for (int k = 0; k < A.num_rows(); ++k) {
        std::string r;
        r.reserve(cols.size());
        for (const auto &c : cols) r += c.first > 0 ? A.row(k)[c.first - 1] : '-';
        out.push_back({(clash ? "A." : "") + A.row_name(k), r});
    }
    for (int l = 0; l < B.num_rows(); ++l) {
        std::string r;
        r.reserve(cols.size());
        for (const auto &c : cols) r += c.second > 0 ? B.row(l)[c.second - 1] : '-';
        out.push_back({(clash ? "B." : "") + B.row_name(l), r});
    }
    return out;
}

std::string clustal_blocks(const std::vector<std::pair<std::string, const std::string *>> &rows, size_t width) {
    size_t namew = 18, len = 0;
    for (const auto &r : rows) {
        namew = std::max(namew, r.first.size());
        len = std::max(len, r.second->size());
    }
    std::string text;
    size_t start = 0;
    do {
        const size_t end = std::min(len, start + width);
        for (const auto &r : rows) {
            text += r.first;
            text.append(namew - r.first.size(), ' ');
            text += ' ';
            text += r.second->substr(std::min(start, r.second->size()), end - start);
            text += '\n';
        }
        start = end;
        if (start < len) text += '\n';
    } while (start < len);
    return text;
}

std::string format_prob(double p) {
    // a stream with precision 4 and the default float field writes what
    // printf's %.4g writes, the scientific field with precision 3 what %.3e
    // writes (the streams' num_put is specified by those conversions)
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.4g", p);
    std::string txt = buf;
    if (txt.size() > 8) {  // longer than precision + 4 characters
        std::snprintf(buf, sizeof buf, "%.3e", p);
        txt = buf;
    }
    const size_t e = txt.find("e-0");
    if (e != std::string::npos) txt.replace(e, 3, "e-");
    return txt;
}

std::string consensus_pp_text(const RnaData &A, const RnaData &B, const std::vector<std::pair<int, int>> &cols,
                              double p_expA, double p_expB, bool stacking) {
    const double wA = A.num_rows(), wB = B.num_rows();
    const double cut = std::exp((std::log(A.cutoff()) * wA + std::log(B.cutoff()) * wB) / (wA + wB));
    const double floor_pen = cut * 0.1;  // penalty factor of rna_data.cc:116-117
    const double floorA = std::min(p_expA, floor_pen), floorB = std::min(p_expB, floor_pen);
    auto mean = [&](double pa, double pb) {
        pa = std::max(floorA, pa);
        pb = std::max(floorB, pb);
        return std::exp((std::log(pa) * wA + std::log(pb) * wB) / (wA + wB));
    };
    // probabilities of the column pairs: only pairs with a base pair on
    // some side can exceed the cut (the floors' mean is below it), so the
    // loop runs over the base pairs of A and of B mapped to columns
    const int ncols = (int)cols.size();
    std::vector<int> colA(A.length() + 1, -1), colB(B.length() + 1, -1);
    for (int c = 0; c < ncols; ++c) {
        if (cols[c].first > 0) colA[cols[c].first] = c;
        if (cols[c].second > 0) colB[cols[c].second] = c;
    }
    std::map<std::pair<int, int>, double> out;
    auto consider = [&](int ci, int cj) {
        if (ci < 0 || cj < 0 || ci >= cj || out.count({ci + 1, cj + 1})) return;
        const auto &ei = cols[ci], &ej = cols[cj];
        const bool okA = ei.first > 0 && ej.first > 0, okB = ei.second > 0 && ej.second > 0;
        const double pc = mean(okA ? A.arc_prob(ei.first, ej.first) : 0, okB ? B.arc_prob(ei.second, ej.second) : 0);
        bool keep = pc > cut;
        if (stacking && !keep) {
            const double sc = mean(okA ? A.joint_arc_prob(ei.first, ej.first) : 0,
                                   okB ? B.joint_arc_prob(ei.second, ej.second) : 0);
            keep = sc > cut;
        }
        if (keep) out[{ci + 1, cj + 1}] = pc;
    };
    for (const auto &x : A.pairs()) consider(colA[x.first.first], colA[x.first.second]);
    for (const auto &x : B.pairs()) consider(colB[x.first.first], colB[x.first.second]);
    std::vector<std::pair<std::string, std::string>> rows = aligned_rows(A, B, cols);
    size_t namew = 18;
    for (const auto &r : rows) namew = std::max(namew, r.first.size());
    std::string text = "#PP 2.0\n\n";
    for (const auto &r : rows) {
        text += r.first;
        text.append(namew - std::min(namew, r.first.size()), ' ');
        text += ' ';
        text += r.second;
        text += '\n';
    }
    text += "\n#END\n\n#SECTION BASEPAIRS\n\n#BPCUT " + format_prob(cut) + "\n\n";
    for (const auto &x : out)
        text += std::to_string(x.first.first) + " " + std::to_string(x.first.second) + " " + format_prob(x.second) + "\n";
    text += "\n#END\n";
    return text;
}

}  // namespace locarna_amd
