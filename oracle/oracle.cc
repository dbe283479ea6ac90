// CPU oracle for the LocARNA pairwise alignment path -- TEST INFRASTRUCTURE.
//
// A direct restatement of the reference algorithm (LocARNA 2.0.1), written
// for clarity rather than speed: 64-bit scores with the reference's
// infinity encoding, one dense M matrix reused by all arc-match groups,
// recursive traceback.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use it; the product (locarna_amd/) never does.
//
// Citations are file:line into the reference source tree (src/LocARNA/...).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <atomic>
#include <queue>
#include <vector>

namespace orc {

typedef long long sc_t;

// infty_int.cc:7-23: finite range [min/5, -min/5); normalized -inf = min/5*2
const sc_t LMIN = std::numeric_limits<sc_t>::min();
const sc_t MIN_FINITE = LMIN / 5;
const sc_t NEG_INF = LMIN / 5 * 2;
inline bool is_neg_inf(sc_t v) { return v < MIN_FINITE; }
inline sc_t normalize(sc_t v) { return v < MIN_FINITE ? NEG_INF : v; }  // infty_int.hh:334-341

struct failure : std::runtime_error {
    explicit failure(const std::string &s) : std::runtime_error(s) {}
};

template <class T>
struct Mat {
    size_t r = 0, c = 0;
    std::vector<T> v;
    void resize(size_t r_, size_t c_, T x) { r = r_; c = c_; v.assign(r * c, x); }
    T &operator()(size_t i, size_t j) { return v[i * c + j]; }
    const T &operator()(size_t i, size_t j) const { return v[i * c + j]; }
};

// ---------------------------------------------------------------- RIBOSUM
// ribosum.cc:42-197, ribosum.hh:384-419 (lower triangular, symmetric fill)
struct Ribosum {
    double bm[4][4];
    double am[16][16];
    double base_nonstruct[4];
    double basepair[4][4];
    double basematch[4][4];
    double arcmatch[16][16];
    static int idx(char c) {
        switch (c) { case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'U': return 3; }
        return -1;
    }
    // ribosum.cc:325-331
    double basematch_score_corrected(char i, char j) const {
        double bi = base_nonstruct[idx(i)], bj = base_nonstruct[idx(j)];
        return log(basematch[idx(i)][idx(j)] / (bi * bj)) / log(2);
    }
};

static void read_tri(std::istream &in, int n, double *m) {
    std::string line;
    while (std::getline(in, line) && line == "") {}
    for (int i = 0; i < n; ++i) {
        if (!std::getline(in, line)) throw failure("ribosum: truncated matrix");
        std::istringstream ls(line);
        std::string name;
        ls >> name;
        for (int j = 0; j <= i; ++j) {
            double x;
            ls >> x;
            m[i * n + j] = m[j * n + i] = x;
        }
    }
}

static void read_block(std::istream &in, const std::string &header, int n, double *m) {
    std::string line;
    auto blank = [](const std::string &s) { return s.find_first_not_of(' ') == std::string::npos; };
    while (std::getline(in, line) && blank(line)) {}
    if (line != header) throw failure("ribosum: expected header " + header);
    for (int i = 0; i < n; ++i) in >> m[i];
}

Ribosum read_ribosum(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw failure("cannot read ribosum file " + path);
    Ribosum r;
    std::string line;
    std::getline(in, line);  // name
    read_tri(in, 4, &r.bm[0][0]);
    std::getline(in, line);
    std::getline(in, line);
    read_tri(in, 16, &r.am[0][0]);
    std::getline(in, line);
    std::getline(in, line);
    double base_probs[4];
    read_block(in, "BASE FREQUENCIES", 4, base_probs);
    read_block(in, "BASE NONSTRUCTURAL FREQUENCIES", 4, r.base_nonstruct);
    read_block(in, "BASE PAIR FREQUENCIES", 16, &r.basepair[0][0]);
    read_block(in, "BASE MATCH FREQUENCIES", 16, &r.basematch[0][0]);
    read_block(in, "BASE PAIR MATCH FREQUENCIES", 256, &r.arcmatch[0][0]);
    return r;
}

// ---------------------------------------------------------------- RnaData
// rna_data.cc:984-1124 (pp 2.0 reader), aux.cc:65-70 (normalize symbols)
struct RnaData {
    std::string name;
    std::string seq;  // the first row; 1-based via seq[i-1]
    // rows of a multiple alignment (a profile), first-seen name order
    // (multiple_alignment.cc:270-380); one row for a single sequence
    std::vector<std::string> names, rows;
    int num_rows() const { return (int)rows.size(); }
    std::map<std::pair<int, int>, double> arc_probs;
    // joint probability of (i, j) and (i+1, j-1) (the #STACK column)
    std::map<std::pair<int, int>, double> arc_2_probs;
    double p_bpcut;
    int length() const { return (int)seq.size(); }
    double arc_prob(int i, int j) const {
        auto it = arc_probs.find({i, j});
        return it == arc_probs.end() ? 0.0 : it->second;
    }
    // rna_data.cc:700-710
    double joint_arc_prob(int i, int j) const {
        auto it = arc_2_probs.find({i, j});
        return it == arc_2_probs.end() ? 0.0 : it->second;
    }
    double stacked_arc_prob(int i, int j) const { return joint_arc_prob(i, j) / arc_prob(i + 1, j - 1); }
    // rna_data.cc:713-732
    double prob_paired_upstream(int i) const {
        double p = 0.0;
        for (int j = i + 1; j <= length(); j++) p += arc_prob(i, j);
        return p;
    }
    double prob_paired_downstream(int i) const {
        double p = 0.0;
        for (int j = 1; j < i; j++) p += arc_prob(j, i);
        return p;
    }
};

static bool get_nonempty_line(std::istream &in, std::string &line) {
    while (std::getline(in, line)) {
        if (line.length() > 0 && !isspace((unsigned char)line[0])) {
            size_t e = line.find_last_not_of(" \t");
            line = line.substr(0, e + 1);
            return true;
        }
    }
    return false;
}

RnaData read_pp(const std::string &text, double min_prob, long long max_bp_span, bool stacking = false) {
    std::istringstream in(text);
    RnaData r;
    r.p_bpcut = min_prob;
    std::string line;
    std::getline(in, line);
    if (line.compare(0, 5, "#PP 2") != 0) throw failure("not a pp 2.0 file");
    // sequence block: "name seq" lines until #END; other '#' lines are annotations
    std::vector<std::string> names;
    std::map<std::string, std::string> seqs;
    while (get_nonempty_line(in, line)) {
        if (line.compare(0, 4, "#END") == 0) break;
        if (line[0] == '#') continue;  // anchor / structure annotation lines
        std::istringstream ls(line);
        std::string name, s;
        ls >> name >> s;
        if (!seqs.count(name)) names.push_back(name);
        seqs[name] += s;
    }
    if (names.empty()) throw failure("no sequence in the pp alignment block");
    for (const std::string &nm : names) {
        std::string row = seqs[nm];
        for (auto &c : row) {  // aux.cc:65-70
            c = (char)toupper((unsigned char)c);
            if (c == 'T') c = 'U';
        }
        if (!r.rows.empty() && row.size() != r.rows[0].size())
            throw failure("rows of the pp alignment differ in length");
        r.rows.push_back(row);
    }
    r.names = names;
    r.name = names[0];
    r.seq = r.rows[0];
    if (!get_nonempty_line(in, line) || line != "#SECTION BASEPAIRS")
        throw failure("Expected base pair section header.");
    size_t max_span = max_bp_span < 0 ? std::numeric_limits<size_t>::max() : (size_t)max_bp_span;
    bool contains_stacking = false;
    while (get_nonempty_line(in, line)) {
        if (line[0] == '#') {
            if (line.compare(0, 4, "#END") == 0) break;
            if (line.compare(0, 6, "#STACK") == 0) contains_stacking = true;
            if (line.compare(0, 6, "#BPCUT") == 0) {
                std::istringstream ls(line);
                std::string dummy;
                double p;
                ls >> dummy >> p;
                if (ls.fail()) throw failure("Cannot parse line \"" + line + "\" in base pairs section.");
                r.p_bpcut = std::max(p, r.p_bpcut);
            }
            continue;
        }
        std::istringstream ls(line);
        size_t i, j;
        double p;
        ls >> i >> j >> p;
        if (ls.fail()) throw failure("Cannot parse line \"" + line + "\" in base pairs section.");
        if (!(1 <= i && i < j && j <= r.seq.size()))
            throw failure("Invalid indices in PP input line \"" + line + "\".");
        if (p <= r.p_bpcut || (j - i + 1) > max_span) continue;  // rna_data.cc:1078
        r.arc_probs[{(int)i, (int)j}] = p;
        // rna_data.cc:1083-1092: with stacking, a fourth column is the joint
        // probability of (i, j) and (i+1, j-1)
        double p2;
        if (stacking && (ls >> p2) && p2 > r.p_bpcut) r.arc_2_probs[{(int)i, (int)j}] = p2;
    }
    if (!contains_stacking && !r.arc_2_probs.empty())  // rna_data.cc:1097-1100
        throw failure("Stacking probabilties found but stack keyword is missing.");
    return r;
}

// ---------------------------------------------------------------- BasePairs
// basepairs.cc:156-220
struct Arc {
    int idx, left, right;
};
struct BasePairs {
    int len;
    std::vector<Arc> arcs;
    std::vector<std::vector<Arc>> left_adj, right_adj;  // with sentinels
    BasePairs(const RnaData &rna, double min_prob) : len(rna.length()) {
        left_adj.resize(len + 1);
        right_adj.resize(len + 1);
        for (int i = len - 3; i >= 1; i--) {
            for (int j = i + 3; j <= len; j++) {
                double p = rna.arc_prob(i, j);
                if (p >= min_prob) {
                    Arc a{(int)arcs.size(), i, j};
                    arcs.push_back(a);
                    left_adj[i].push_back(a);
                    right_adj[j].push_back(a);
                }
            }
        }
        for (auto &x : left_adj)
            std::sort(x.begin(), x.end(), [](const Arc &a, const Arc &b) { return a.right < b.right; });
        for (auto &x : right_adj)
            std::sort(x.begin(), x.end(), [](const Arc &a, const Arc &b) { return a.left > b.left; });
        for (auto &x : left_adj) x.push_back(Arc{0, len + 1, len + 1});
        for (auto &x : right_adj) x.push_back(Arc{0, 0, 0});
    }
};

// ---------------------------------------------------------------- band
// trace_controller.hh:328-336; trace_controller.cc:406-440 (max-diff without
// a reference alignment), :565-597 (restrict by trace probabilities)
struct Band {
    int lenA, lenB;
    std::vector<size_t> min_col, max_col;
    Band(int a, int b) : lenA(a), lenB(b), min_col(a + 1, 0), max_col(a + 1, (size_t)b) {}
    // trace_controller.cc:361-390, unsigned arithmetic as there
    void constrain_wo_ref(size_t delta) {
        const size_t la = lenA, lb = lenB;
        if (la == 0 || lb == 0) return;
        for (size_t i = 0; i <= la; i++) {
            size_t x = i * lb * (la + lb);
            size_t y = 2 * delta * la * lb;
            size_t z = la * (la + lb);
            if (la > lb)
                y = std::max(y, (la + lb) * la / 2);
            else if (lb > la)
                y = std::max(y, (la + lb) * lb / 2);
            min_col[i] = x > y ? (x - y + z - 1) / z : 0;
            max_col[i] = std::min((x + y) / z, lb);
        }
    }
    bool is_valid(size_t i, size_t j) const { return min_col[i] <= j && j <= max_col[i]; }
    bool is_valid_match(size_t i, size_t j) const { return is_valid(i, j) && is_valid(i - 1, j - 1); }
    Band reverse() const {  // trace_controller.cc:320-338
        Band r(lenA, lenB);
        for (int i = 0; i <= lenA; ++i) {
            r.max_col[lenA - i] = lenB - min_col[i];
            r.min_col[lenA - i] = lenB - max_col[i];
        }
        return r;
    }
};

// stral_score.cc:10-98
struct StralScore {
    std::vector<std::string> seqA, seqB;  // rows
    const double (*sim)[4];
    double struct_weight, indel_opening, indel;
    std::vector<double> p_upA, p_downA, p_upB, p_downB;
    StralScore(const RnaData &a, const RnaData &b, const double (*sim_)[4], double sw, double io, double ind)
        : seqA(a.rows), seqB(b.rows), sim(sim_), struct_weight(sw), indel_opening(io), indel(ind) {
        init(a, p_upA, p_downA);
        init(b, p_upB, p_downB);
    }
    static void init(const RnaData &r, std::vector<double> &up, std::vector<double> &down) {
        int n = r.length();
        up.assign(n + 1, 0.0);
        down.assign(n + 1, 0.0);
        for (int i = 1; i <= n; i++) {
            up[i] = r.prob_paired_upstream(i);
            down[i] = r.prob_paired_downstream(i);
        }
    }
    double sigma(size_t i, size_t j) const {
        // average over the row pairs whose symbols are both in ACGU
        int pairs = 0;
        double seq_score = 0;
        for (size_t k = 0; k < seqA.size(); k++)
            for (size_t l = 0; l < seqB.size(); l++) {
                int a = Ribosum::idx(seqA[k][i - 1]), b = Ribosum::idx(seqB[l][j - 1]);
                if (a >= 0 && b >= 0) {
                    seq_score += sim[a][b];
                    pairs++;
                }
            }
        if (pairs != 0) seq_score /= pairs;
        return struct_weight * (sqrt(p_downA[i] * p_downB[j]) + sqrt(p_upA[i] * p_upB[j])) + seq_score;
    }
    void reverse() {
        for (auto &r : seqA) std::reverse(r.begin(), r.end());
        for (auto &r : seqB) std::reverse(r.begin(), r.end());
        std::reverse(p_upA.begin() + 1, p_upA.end());
        std::reverse(p_downA.begin() + 1, p_downA.end());
        std::reverse(p_upB.begin() + 1, p_upB.end());
        std::reverse(p_downB.begin() + 1, p_downB.end());
        std::swap(p_upA, p_downA);
        std::swap(p_upB, p_downB);
    }
};

// edge_probs.icc:7-162 (PFGotoh), :212-268 (PFTraceProbs); flag_local =
// sequ_local (locarna's band never uses the local variant for global
// alignment).  Free end gaps (free_endgaps.hh, bits as oracle_params): the
// forward pass adds 1 to A(i, 0) with the second sequence's left flag and to
// B(0, j) with the first's (edge_probs.icc:120-131), the reverse pass does
// the same with the flags of the reversed sequences (left and right
// swapped, FreeEndgaps::reverse), and the global Z adds the A entries of the
// last column and the B entries of the last row under the same left flags
// of the forward pass (:39-53).
inline bool fe_left_1(int f) { return f & 1; }
inline bool fe_right_1(int f) { return f & 2; }
inline bool fe_left_2(int f) { return f & 4; }
inline bool fe_right_2(int f) { return f & 8; }
inline int fe_reverse(int f) {  // {right_1, left_1, right_2, left_2}
    return (fe_right_1(f) ? 1 : 0) | (fe_left_1(f) ? 2 : 0) | (fe_right_2(f) ? 4 : 0) | (fe_left_2(f) ? 8 : 0);
}

template <class pf_t>
struct PFGotoh {
    int lenA, lenB;
    double temp;
    bool local;
    int fe;
    Mat<pf_t> zM, zA, zB, zMr, zAr, zBr;
    pf_t z;
    void pf_gotoh(Mat<pf_t> &zMl, Mat<pf_t> &zAl, Mat<pf_t> &zBl, const Band &tc, const StralScore &score, int f) {
        double g_open = exp(score.indel_opening / temp);
        double g_ext = exp(score.indel / temp);
        zMl.resize(lenA + 1, lenB + 1, 0);
        zAl.resize(lenA + 1, lenB + 1, 0);
        zBl.resize(lenA + 1, lenB + 1, 0);
        if (tc.is_valid(0, 0)) zMl(0, 0) = local ? 0 : 1;
        if (lenA > 0 && tc.is_valid(1, 0)) zAl(1, 0) = g_open * g_ext;
        if (lenB > 0 && tc.is_valid(0, 1)) zBl(0, 1) = g_open * g_ext;
        for (int i = 2; i <= lenA; i++) {
            if (tc.min_col[i] > 0) break;
            zAl(i, 0) = zAl(i - 1, 0) * g_ext;
        }
        for (size_t j = std::max(tc.min_col[0], (size_t)2); j <= std::min(tc.max_col[0], (size_t)lenB); j++)
            zBl(0, j) = zBl(0, j - 1) * g_ext;
        if (fe_left_2(f))
            for (int i = 1; i <= lenA; i++) zAl(i, 0) += 1;
        if (fe_left_1(f))
            for (int j = 1; j <= lenB; j++) zBl(0, j) += 1;
        for (int i = 1; i <= lenA; i++) {
            for (size_t j = std::max(tc.min_col[i], (size_t)1); j <= std::min(tc.max_col[i], (size_t)lenB); j++) {
                double match_score_ij = score.sigma(i, j);
                double match_ij = exp(match_score_ij / temp);
                zMl(i, j) = zMl(i - 1, j - 1) * match_ij + zAl(i - 1, j - 1) * match_ij +
                            zBl(i - 1, j - 1) * match_ij + (local ? match_ij : 0);
                zAl(i, j) = zAl(i - 1, j) * g_ext + zMl(i - 1, j) * g_open * g_ext + zBl(i - 1, j) * g_open * g_ext;
                zBl(i, j) = zBl(i, j - 1) * g_ext + zMl(i, j - 1) * g_open * g_ext + zAl(i, j - 1) * g_open * g_ext;
            }
        }
    }
    PFGotoh(const RnaData &a, const RnaData &b, const Band &tc, const double (*sim)[4], double gap_opening,
            double gap_extension, double sw, double temp_, bool local_, int fe_)
        : lenA(a.length()), lenB(b.length()), temp(temp_), local(local_), fe(fe_) {
        StralScore score(a, b, sim, sw, gap_opening, gap_extension);
        pf_gotoh(zM, zA, zB, tc, score, fe);
        score.reverse();
        pf_gotoh(zMr, zAr, zBr, tc.reverse(), score, fe_reverse(fe));
        if (local) {
            z = 1;
            for (int i = 0; i <= lenA; i++)
                for (int j = 0; j <= lenB; j++) z += zM(i, j);
        } else {
            z = zM(lenA, lenB) + zA(lenA, lenB) + zB(lenA, lenB);
            if (fe_left_2(fe))
                for (int i = 0; i <= lenA; i++) z += zA(i, lenB);
            if (fe_left_1(fe))
                for (int j = 0; j <= lenB; j++) z += zB(lenA, j);
        }
    }
    bool fail() const { return z == (pf_t)0 || std::isnan((double)z) || std::isinf((double)z); }
};

template <class pf_t>
Mat<double> trace_probs(const RnaData &a, const RnaData &b, const Band &tc, const double (*sim)[4],
                        double gap_opening, double gap_extension, double sw, double temp, bool local, int fe) {
    PFGotoh<pf_t> g(a, b, tc, sim, gap_opening, gap_extension, sw, temp, local, fe);
    Mat<double> probs;
    int lenA = a.length(), lenB = b.length();
    probs.resize(lenA + 1, lenB + 1, 0.0);
    double locality_add = local ? 1 : 0;
    double g_open = exp(gap_opening / temp);
    for (int i = 0; i <= lenA; i++) {
        for (size_t j = std::max(tc.min_col[i], (size_t)0); j <= std::min(tc.max_col[i], (size_t)lenB); j++) {
            size_t ri = lenA - i, rj = lenB - j;
            pf_t z_ij = g.zM(i, j) * (g.zMr(ri, rj) + g.zAr(ri, rj) + g.zBr(ri, rj) + locality_add) +
                        g.zA(i, j) * (g.zMr(ri, rj) + g.zAr(ri, rj) / g_open + g.zBr(ri, rj)) +
                        g.zB(i, j) * (g.zMr(ri, rj) + g.zAr(ri, rj) + g.zBr(ri, rj) / g_open);
            probs(i, j) = z_ij / g.z;
        }
    }
    return probs;
}

// trace_controller.cc:565-597
void restrict_by_trace_probabilities(Band &tc, const Mat<double> &probs, double min_prob) {
    int rows = tc.lenA;
    for (int i = 0; i <= rows; ++i) {
        size_t new_min = tc.max_col[i];
        size_t new_max = tc.min_col[i];
        for (size_t j = tc.min_col[i]; j <= tc.max_col[i]; ++j) {
            if (probs(i, j) >= min_prob) {
                new_min = std::min(new_min, j);
                new_max = std::max(new_max, j);
            }
        }
        tc.min_col[i] = std::max(tc.min_col[i], new_min);
        tc.max_col[i] = std::min(tc.max_col[i], new_max);
    }
    size_t max_col = 0;
    for (int i = 0; i <= rows; ++i) {
        tc.max_col[i] = std::max(tc.max_col[i], max_col);
        max_col = tc.max_col[i];
    }
    size_t min_col = tc.max_col[rows];
    for (int i = rows + 1; i > 0;) {
        --i;
        tc.min_col[i] = std::min(tc.min_col[i], min_col);
        min_col = tc.min_col[i];
    }
}

// ---------------------------------------------------------------- ArcMatches
// arc_matches.cc:19-48, 130-188, 76-110, 313-355
struct ArcMatch {
    int idx;
    Arc a, b;
};
struct ArcMatches {
    int lenA, lenB;
    std::vector<ArcMatch> ams;
    Mat<std::vector<int>> left_lists, right_lists;
    ArcMatches(const BasePairs &A, const BasePairs &B, const Band &tc, int diff_am = -1, int diff_at_am = -1)
        : lenA(A.len), lenB(B.len) {
        left_lists.resize(lenA + 1, lenB + 1, {});
        right_lists.resize(lenA + 1, lenB + 1, {});
        // locarna.cc:606-611: -1 means the longer length
        size_t max_length_diff = diff_am != -1 ? (size_t)diff_am : (size_t)std::max(lenA, lenB);
        size_t max_diff_at_am = diff_at_am != -1 ? (size_t)diff_at_am : (size_t)std::max(lenA, lenB);
        for (const Arc &a : A.arcs) {
            for (const Arc &b : B.arcs) {
                bool valid = tc.is_valid_match(a.left, b.left) && tc.is_valid_match(a.right, b.right) &&
                             ((size_t)abs((a.right - a.left) - (b.right - b.left)) <= max_length_diff) &&
                             (size_t)a.left <= b.left + max_diff_at_am && (size_t)b.left <= a.left + max_diff_at_am &&
                             (size_t)a.right <= b.right + max_diff_at_am && (size_t)b.right <= a.right + max_diff_at_am;
                if (!valid) continue;
                int idx = (int)ams.size();
                ams.push_back(ArcMatch{idx, a, b});
                left_lists(a.left, b.left).push_back(idx);
                right_lists(a.right, b.right).push_back(idx);
            }
        }
        for (int i = 1; i <= lenA; i++)
            for (int j = 1; j <= lenB; j++) {
                auto &l = right_lists(i, j);
                std::sort(l.begin(), l.end(), [&](int x, int y) {
                    const ArcMatch &u = ams[x], &v = ams[y];
                    return u.a.left > v.a.left || (u.a.left == v.a.left && u.b.left > v.b.left);
                });
            }
    }
    void max_right_ends(int al, int bl, int *max_ar, int *max_br) const {
        for (int x : left_lists(al, bl)) {
            *max_ar = std::max(*max_ar, ams[x].a.right);
            *max_br = std::max(*max_br, ams[x].b.right);
        }
    }
    // arc_matches.cc:50-74: the arc match of the arcs directly inside both
    // arcs (left ends + 1, right ends - 1), the first in its left-end list;
    // -1 when there is none
    std::vector<int> inner;
    void init_inner() {
        inner.assign(ams.size(), -1);
        for (const ArcMatch &am : ams) {
            if (am.a.left + 1 > lenA || am.b.left + 1 > lenB) continue;
            for (int y : left_lists(am.a.left + 1, am.b.left + 1))
                if (ams[y].a.right == am.a.right - 1 && ams[y].b.right == am.b.right - 1) {
                    inner[am.idx] = y;
                    break;
                }
        }
    }
    // arc_matches.cc:313-355 with no_lonely_pairs: the group (al, bl) holds
    // the inner region of the arc matches with left ends (al-1, bl-1) that
    // have an inner arc match
    void max_right_ends_noLP(int al, int bl, int *max_ar, int *max_br) const {
        int ar = *max_ar + 1, br = *max_br + 1;
        for (int x : left_lists(al - 1, bl - 1)) {
            if (inner[x] < 0) continue;
            ar = std::max(ar, ams[x].a.right);
            br = std::max(br, ams[x].b.right);
        }
        *max_ar = ar - 1;
        *max_br = br - 1;
    }
};

// ---------------------------------------------------------------- Scoring
// scoring.cc:28-52, 140-198 (sigma), 200-265 (weights), 271-311 (gap),
// 368-438 (ribosum arc-match term), 440-554 (arcmatch)
struct Scoring {
    const oracle_params &p;
    const Ribosum *rib;
    const RnaData &A, &B;
    const BasePairs &bpsA, &bpsB;
    Mat<sc_t> sigma_tab;
    std::vector<sc_t> gapA_tab, gapB_tab, weightsA, weightsB;
    std::vector<sc_t> stack_weightsA, stack_weightsB;  // with --stacking / --new-stacking
    static sc_t round2score(double d) { return (sc_t)((d < 0) ? (d - 0.5) : (d + 0.5)); }  // scoring.hh:384-387
    Scoring(const oracle_params &p_, const Ribosum *r, const RnaData &a, const RnaData &b, const BasePairs &ba,
            const BasePairs &bb)
        : p(p_), rib(r), A(a), B(b), bpsA(ba), bpsB(bb) {
        int lenA = A.length(), lenB = B.length();
        sigma_tab.resize(lenA + 1, lenB + 1, 0);
        for (int i = 1; i <= lenA; ++i)
            for (int j = 1; j <= lenB; ++j) sigma_tab(i, j) = sigma_(i, j);
        gapA_tab.assign(lenA + 1, 0);
        gapB_tab.assign(lenB + 1, 0);
        // scoring.cc:271-311: gap frequencies of the columns in float, the
        // gap cost round((1 - freq) * indel)
        auto gapcost = [&](const RnaData &r, std::vector<sc_t> &tab) {
            for (int i = 1; i <= r.length(); i++) {
                float freq = 0;
                for (const std::string &row : r.rows) freq += (row[i - 1] == '-') ? 1 : 0;
                freq /= r.rows.size();
                tab[i] = round2score((1 - freq) * p.indel);
            }
        };
        gapcost(A, gapA_tab);
        gapcost(B, gapB_tab);
        double peA = p.exp_prob > 0 ? p.exp_prob : 1.0 / (2.0 * lenA);  // aux.hh:216-220
        double peB = p.exp_prob > 0 ? p.exp_prob : 1.0 / (2.0 * lenB);
        weights(A, bpsA, peA, weightsA);
        weights(B, bpsB, peB, weightsB);
        stack_weights(A, bpsA, peA, weightsA, stack_weightsA);
        stack_weights(B, bpsB, peB, weightsB, stack_weightsB);
    }
    bool stacking() const { return p.stacking != 0; }
    // sum of pairs over the rows of the two columns, then the integer
    // quotient by the number of row pairs (C++ division, toward zero)
    sc_t sigma_(int ia, int ib) const {
        sc_t score = 0;
        for (const std::string &ra : A.rows)
            for (const std::string &rb : B.rows) {
                char cA = ra[ia - 1], cB = rb[ib - 1];
                if (rib && Ribosum::idx(cA) >= 0 && Ribosum::idx(cB) >= 0) {
                    score += round2score(100.0 * rib->basematch_score_corrected(cA, cB));
                } else if (cA != 'N' && cB != 'N') {
                    score += (cA == cB) ? p.match : p.mismatch;
                }
            }
        return round2score(score / (int)(A.rows.size() * B.rows.size()));
    }
    sc_t probToWeight(double prob, double pe) const {
        return round2score(round((sc_t)p.struct_weight * (1 - log(prob) / log(pe))));
    }
    void weights(const RnaData &r, const BasePairs &bps, double pe, std::vector<sc_t> &w) {
        w.resize(bps.arcs.size());
        for (size_t i = 0; i < bps.arcs.size(); ++i)
            w[i] = probToWeight(r.arc_prob(bps.arcs[i].left, bps.arcs[i].right), pe);
    }
    // scoring.cc:200-236: the stacked weight of an arc whose inner arc
    // exists (probability > 0); others keep 0.  Only arc matches of
    // stackable arcs (joint probability > 0 in both) use these weights, and
    // a stackable arc's joint probability exceeds the cut, so the logarithm
    // is finite wherever it is used.
    void stack_weights(const RnaData &r, const BasePairs &bps, double pe, const std::vector<sc_t> &w,
                       std::vector<sc_t> &sw) {
        sw.assign(bps.arcs.size(), 0);
        const bool old_st = (p.stacking & 1) != 0, new_st = (p.stacking & 2) != 0;
        for (size_t i = 0; i < bps.arcs.size(); ++i) {
            const Arc &a = bps.arcs[i];
            const bool inner = r.arc_prob(a.left + 1, a.right - 1) > 0;
            const bool joint = r.joint_arc_prob(a.left, a.right) > 0;
            if (old_st && inner && joint) sw[i] = probToWeight(r.stacked_arc_prob(a.left, a.right), pe);
            if (new_st) {
                if (!old_st) sw[i] = w[i];
                if (inner && joint) sw[i] += probToWeight(r.joint_arc_prob(a.left, a.right), pe);
            }
        }
    }
    // scoring.cc:556-569
    bool stackable(const ArcMatch &am) const {
        return A.joint_arc_prob(am.a.left, am.a.right) > 0 && B.joint_arc_prob(am.b.left, am.b.right) > 0;
    }
    // the log-odds averaged over the row pairs whose four end symbols are
    // all in ACGU (rows with a gap at an end take no part)
    sc_t riboX_arcmatch_score(const Arc &a, const Arc &b) const {
        double score = 0;
        int considered = 0;
        for (const std::string &ra : A.rows)
            for (const std::string &rb : B.rows) {
                char al = ra[a.left - 1], ar = ra[a.right - 1], bl = rb[b.left - 1], br = rb[b.right - 1];
                if (Ribosum::idx(al) >= 0 && Ribosum::idx(ar) >= 0 && Ribosum::idx(bl) >= 0 &&
                    Ribosum::idx(br) >= 0) {
                    considered++;
                    double pam = rib->arcmatch[Ribosum::idx(al) * 4 + Ribosum::idx(ar)]
                                              [Ribosum::idx(bl) * 4 + Ribosum::idx(br)];
                    double pa = rib->basepair[Ribosum::idx(al)][Ribosum::idx(ar)];
                    double pb = rib->basepair[Ribosum::idx(bl)][Ribosum::idx(br)];
                    score += log(pam / (pa * pb)) / log(2);
                }
            }
        if (considered == 0) return 0;
        return round2score(100.0 * score / considered);
    }
    sc_t arcmatch(const ArcMatch &am, bool stacked = false) const {
        sc_t seq_contrib = 0;
        if (p.tau != 0) {
            if (rib) seq_contrib = riboX_arcmatch_score(am.a, am.b);
            else seq_contrib = sigma_tab(am.a.left, am.b.left) + sigma_tab(am.a.right, am.b.right);
        }
        return ((sc_t)p.tau * seq_contrib) / 100 +
               (stacked ? stack_weightsA[am.a.idx] + stack_weightsB[am.b.idx] : weightsA[am.a.idx] + weightsB[am.b.idx]);
    }
};

// ------------------------------------------------------------- consensus
// The pp file of an aligned pair, as locarna --pp writes it for the next
// step of mlocarna's progressive alignment: rows from the alignment edges
// (multiple_alignment.cc:186-247, "A." / "B." name prefixes on a name
// clash), consensus base-pair probabilities as weighted geometric means
// (rna_data.cc:1475-1570, penalty factor 0.1 from :116-120), written by
// RnaData::write_pp (rna_data.cc:1242-1339, probabilities by format_prob,
// name / sequence lines by multiple_alignment.cc:995-1030).  The reference
// iterates a hash map for the base-pair lines; they are written here in
// (i, j) order, which the reader does not see.
std::string format_prob(double prob) {
    const int precision = 4;
    std::ostringstream outd;
    outd.precision(precision);
    outd << prob;
    std::string str = outd.str();
    if (str.length() > (size_t)precision + 4) {
        std::ostringstream outs;
        outs.setf(std::ios::scientific, std::ios::floatfield);
        outs.precision(precision - 1);
        outs << prob;
        str = outs.str();
    }
    size_t pos = str.find("e-0");
    if (pos != std::string::npos) str.replace(pos, 3, "e-");
    return str;
}

std::string consensus_pp(const RnaData &A, const RnaData &B, const std::vector<std::pair<int, int>> &edges,
                         double p_expA, double p_expB, bool stacking) {
    const size_t rowsA = A.rows.size(), rowsB = B.rows.size();
    std::vector<std::string> aliA(rowsA), aliB(rowsB);
    for (const auto &e : edges) {
        for (size_t k = 0; k < rowsA; ++k) aliA[k] += e.first < 0 ? '-' : A.rows[k][e.first - 1];
        for (size_t k = 0; k < rowsB; ++k) aliB[k] += e.second < 0 ? '-' : B.rows[k][e.second - 1];
    }
    bool clash = false;
    for (const std::string &x : A.names)
        for (const std::string &y : B.names) clash = clash || x == y;
    std::vector<std::pair<std::string, std::string>> rows;
    for (size_t k = 0; k < rowsA; ++k) rows.push_back({(clash ? "A." : "") + A.names[k], aliA[k]});
    for (size_t k = 0; k < rowsB; ++k) rows.push_back({(clash ? "B." : "") + B.names[k], aliB[k]});
    const double p_minMean = exp((log(A.p_bpcut) * rowsA + log(B.p_bpcut) * rowsB) / (rowsA + rowsB));
    const double p_penalty = p_minMean * 0.1;
    auto cons = [&](double pA, double pB) {
        pA = std::max(std::min(p_expA, p_penalty), pA);
        pB = std::max(std::min(p_expB, p_penalty), pB);
        return exp((log(pA) * rowsA + log(pB) * rowsB) / (rowsA + rowsB));
    };
    std::map<std::pair<int, int>, double> probs;
    for (size_t i = 0; i < edges.size(); i++)
        for (size_t j = i + 1; j < edges.size(); j++) {
            const auto &ei = edges[i], &ej = edges[j];
            const double pA = (ei.first < 0 || ej.first < 0) ? 0 : A.arc_prob(ei.first, ej.first);
            const double pB = (ei.second < 0 || ej.second < 0) ? 0 : B.arc_prob(ei.second, ej.second);
            const double pc = cons(pA, pB);
            if (stacking) {
                const double sA = (ei.first < 0 || ej.first < 0) ? 0 : A.joint_arc_prob(ei.first, ej.first);
                const double sB = (ei.second < 0 || ej.second < 0) ? 0 : B.joint_arc_prob(ei.second, ej.second);
                // the joint consensus decides membership; the written file
                // carries no #STACK column (the consensus object is built
                // with has_stacking_ false)
                if (pc > p_minMean || cons(sA, sB) > p_minMean) probs[{(int)i + 1, (int)j + 1}] = pc;
            } else if (pc > p_minMean) {
                probs[{(int)i + 1, (int)j + 1}] = pc;
            }
        }
    size_t namewidth = 18;
    for (const auto &r : rows) namewidth = std::max(namewidth, r.first.length());
    std::ostringstream out;
    out << "#PP 2.0" << std::endl << std::endl;
    for (const auto &r : rows) {
        std::streamsize ow = out.width(namewidth);
        out << std::left << r.first << " ";
        out.width(ow);
        out << r.second << std::endl;
    }
    out << std::endl << "#END" << std::endl;
    out << std::endl << "#SECTION BASEPAIRS" << std::endl << std::endl << "#BPCUT " << format_prob(p_minMean)
        << std::endl << std::endl;
    for (const auto &x : probs) out << x.first.first << " " << x.first.second << " " << format_prob(x.second) << std::endl;
    out << std::endl << "#END" << std::endl;
    return out.str();
}

// ---------------------------------------------------------------- inputs
struct Inputs {
    oracle_params p;
    std::unique_ptr<Ribosum> rib;
    RnaData A, B;
    std::unique_ptr<BasePairs> bpsA, bpsB;
    std::unique_ptr<Band> band;
    std::unique_ptr<ArcMatches> am;
    std::unique_ptr<Scoring> sc;
};

// src/locarna.cc:398-703 and src/locarna_p.cc:300-477 (parameter plumbing)
std::unique_ptr<Inputs> prepare(const char *ppA, const char *ppB, const oracle_params &p, const char *rib_path) {
    if (p.sequ_local && p.free_endgaps)  // locarna.cc:375-378
        throw std::runtime_error("Free endgaps cannot be combined with local alignment.");
    if (p.struct_local && (p.no_lonely_pairs || p.stacking))  // restated without these combinations
        throw std::runtime_error("--struct-local with --noLP or stacking is not restated");
    auto in = std::make_unique<Inputs>();
    in->p = p;
    if (p.use_ribosum) in->rib.reset(new Ribosum(read_ribosum(rib_path)));
    in->A = read_pp(ppA, p.min_prob, p.max_bp_span, p.stacking != 0);
    in->B = read_pp(ppB, p.min_prob, p.max_bp_span, p.stacking != 0);
    int lenA = in->A.length(), lenB = in->B.length();
    in->band.reset(new Band(lenA, lenB));
    if (p.max_diff >= 0) in->band->constrain_wo_ref((size_t)p.max_diff);  // locarna.cc:568-569
    if (p.min_trace_prob > 0.0) {
        // main_helper.icc:371-426: basematch scores = ribosum bm matrix (or
        // match/mismatch); gap costs, struct weight, temperature divided by 100
        double simm[4][4];
        if (in->rib) {
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) simm[i][j] = in->rib->bm[i][j];
        } else {
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) simm[i][j] = i == j ? p.match : p.mismatch;
        }
        Mat<double> probs;
        if (p.band_extended)
            probs = trace_probs<long double>(in->A, in->B, *in->band, simm, p.indel_opening / 100.0, p.indel / 100.0,
                                             p.struct_weight / 100.0, p.temperature_alipf / 100.0, p.sequ_local, p.free_endgaps);
        else
            probs = trace_probs<double>(in->A, in->B, *in->band, simm, p.indel_opening / 100.0, p.indel / 100.0,
                                        p.struct_weight / 100.0, p.temperature_alipf / 100.0, p.sequ_local, p.free_endgaps);
        restrict_by_trace_probabilities(*in->band, probs, p.min_trace_prob);
    }
    in->bpsA.reset(new BasePairs(in->A, p.min_prob));
    in->bpsB.reset(new BasePairs(in->B, p.min_prob));
    in->am.reset(new ArcMatches(*in->bpsA, *in->bpsB, *in->band, p.max_diff_am, p.max_diff_at_am));
    in->am->init_inner();
    in->sc.reset(new Scoring(p, in->rib.get(), in->A, in->B, *in->bpsA, *in->bpsB));
    return in;
}

// ---------------------------------------------------------------- Aligner
// aligner.cc (state E_NO_NO only: no structure locality, no anchors, no
// stacking; --noLP D-fill and traceback; free end gaps at the top level)
struct Aligner {
    Inputs &in;
    const Band &tc;
    const ArcMatches &ams;
    const Scoring &sc;
    const BasePairs &bpsA, &bpsB;
    int lenA, lenB;
    sc_t alpha;
    Mat<sc_t> M;
    std::vector<sc_t> E;
    sc_t F;
    Mat<sc_t> Dmat;
    std::vector<sc_t> am_score;
    std::vector<sc_t> am_score_st;  // stacked score (--stacking)
    std::vector<char> am_stackable;
    int max_i = 0, max_j = 0;
    int min_i = 1, min_j = 1;  // aligner.cc:111-112, set where a local trace stops
    // position penalty of the top level (ModifiedScoringView,
    // aligner_impl.hh:197-262; Scoring::modify_by_parameter, scoring.cc:77-90):
    // sigma - 2 lam, gap costs - lam, D(am) - lam (arc lengths of A and B)
    sc_t lam = 0;
    // AlignerRestriction (aligner_restriction.hh:26-120): the top level
    // aligns A[rsA..reA] with B[rsB..reB] (the whole sequences by default)
    int rsA = 1, reA = 0, rsB = 1, reB = 0;
    bool D_done = false;
    std::vector<std::pair<int, int>> edges;
    std::string structA, structB;
    long long cnt_G = 0, cnt_C = 0, cnt_K = 0;
    bool counting = false;
    // --struct-local (aligner.cc:118-128; states aligner_impl.hh:108-122):
    // E = exclusion, NO = none, X = one (closed), OP = open; the first part
    // refers to A, the second to B.  State 0 is M / E / F above.
    enum { E_NO_NO, E_X_NO, E_NO_X, E_X_X, E_OP_NO, E_NO_OP, E_OP_X, E_X_OP };
    Mat<sc_t> Mx[8];
    std::vector<sc_t> Ex[4];
    sc_t Fx[4] = {0, 0, 0, 0};
    Mat<sc_t> &Ms(int k) { return k ? Mx[k] : M; }
    std::vector<sc_t> &Es(int k) { return k ? Ex[k] : E; }
    sc_t &Fs(int k) { return k ? Fx[k] : F; }
    bool struct_local() const { return in.p.struct_local != 0; }

    explicit Aligner(Inputs &in_)
        : in(in_), tc(*in_.band), ams(*in_.am), sc(*in_.sc), bpsA(*in_.bpsA), bpsB(*in_.bpsB),
          lenA(in_.A.length()), lenB(in_.B.length()), alpha(in_.p.indel_opening) {
        M.resize(lenA + 1, lenB + 1, 0);
        E.assign(lenB + 1, 0);
        if (in_.p.struct_local) {
            for (int k = 1; k < 8; ++k) Mx[k].resize(lenA + 1, lenB + 1, 0);
            for (int k = 1; k < 4; ++k) Ex[k].assign(lenB + 1, 0);
        }
        Dmat.resize(bpsA.arcs.size(), bpsB.arcs.size(), NEG_INF);  // aligner.cc:122-123
        am_score.resize(ams.ams.size());
        for (size_t k = 0; k < ams.ams.size(); ++k) am_score[k] = sc.arcmatch(ams.ams[k]);
        am_score_st.assign(ams.ams.size(), 0);
        am_stackable.assign(ams.ams.size(), 0);
        if (sc.stacking())
            for (size_t k = 0; k < ams.ams.size(); ++k) {
                am_stackable[k] = sc.stackable(ams.ams[k]);
                if (am_stackable[k]) am_score_st[k] = sc.arcmatch(ams.ams[k], true);
            }
        structA.assign(lenA, '.');
        structB.assign(lenB, '.');
        reA = lenA;
        reB = lenB;
    }
    sc_t D(const Arc &a, const Arc &b) const { return Dmat(a.idx, b.idx); }
    sc_t &D(const ArcMatch &am) { return Dmat(am.a.idx, am.b.idx); }

    // aligner.cc:153-233
    sc_t align_noex(size_t al, size_t bl, size_t i, size_t j, sc_t lm = 0) {
        sc_t max_score = M(i - 1, j - 1) + (sc.sigma_tab(i, j) - 2 * lm);
        E[j] = std::max(E[j] + (sc.gapA_tab[i] - lm), M(i - 1, j) + (sc.gapA_tab[i] - lm) + alpha);
        max_score = std::max(max_score, E[j]);
        F = std::max(F + (sc.gapB_tab[j] - lm), M(i, j - 1) + (sc.gapB_tab[j] - lm) + alpha);
        max_score = std::max(max_score, F);
        const auto &adjlA = bpsA.right_adj[i];
        const auto &adjlB = bpsB.right_adj[j];
        for (auto arcA = adjlA.begin(); (size_t)arcA->left > al; ++arcA) {
            for (auto arcB = adjlB.begin(); (size_t)arcB->left > bl; ++arcB) {
                sc_t ns = M(arcA->left - 1, arcB->left - 1) + D(*arcA, *arcB) -
                          lm * (sc_t)((arcA->right - arcA->left + 1) + (arcB->right - arcB->left + 1));
                if (ns <= max_score) continue;
                max_score = ns;
            }
        }
        if (counting) {
            cnt_C++;
            for (int x : ams.right_lists(i, j))
                if ((size_t)ams.ams[x].a.left > al && (size_t)ams.ams[x].b.left > bl) cnt_K++;
        }
        return normalize(max_score);
    }

    // aligner.cc:264-368 (exclA = exclB = false)
    void init_state(size_t al, size_t ar, size_t bl, size_t br, bool globalA, bool globalB, sc_t lm = 0) {
        for (size_t j = bl; j < br; j++) E[j] = NEG_INF;
        M(al, bl) = 0;
        sc_t indel_score = globalA ? alpha : 0;
        size_t i;
        for (i = al + 1; i < ar; i++) {
            if (tc.min_col[i] > bl) break;
            if (!is_neg_inf(indel_score) && globalA) indel_score += sc.gapA_tab[i] - lm;
            M(i, bl) = indel_score;
        }
        for (; i < ar; i++) M(i, tc.min_col[i] - 1) = NEG_INF;
        indel_score = globalB ? alpha : 0;
        size_t j;
        for (j = bl + 1; j < std::min(br, tc.max_col[al] + 1); j++) {
            if (!is_neg_inf(indel_score) && globalB) indel_score += sc.gapB_tab[j] - lm;
            M(al, j) = indel_score;
        }
        for (i = al + 1; i < ar; i++)
            for (; j < std::min(br, tc.max_col[i] + 1); ++j) M(i - 1, j) = NEG_INF;
    }

    // aligner.cc:373-437
    void align_in_arcmatch(size_t al, size_t ar, size_t bl, size_t br) {
        init_state(al, ar, bl, br, true, true);
        for (size_t i = al + 1; i < ar; i++) {
            F = NEG_INF;
            size_t min_col = std::max(bl + 1, tc.min_col[i]);
            size_t max_col = std::min(br - 1, tc.max_col[i]);
            for (size_t j = min_col; j <= max_col; j++) M(i, j) = align_noex(al, bl, i, j);
        }
    }

    // aligner.cc:153-233 in state k < 4 (--struct-local), no position penalty
    sc_t align_noex_s(int k, size_t al, size_t bl, size_t i, size_t j) {
        Mat<sc_t> &Mk = Ms(k);
        std::vector<sc_t> &Ek = Es(k);
        sc_t &Fk = Fs(k);
        sc_t max_score = Mk(i - 1, j - 1) + sc.sigma_tab(i, j);
        Ek[j] = std::max(Ek[j] + sc.gapA_tab[i], Mk(i - 1, j) + sc.gapA_tab[i] + alpha);
        max_score = std::max(max_score, Ek[j]);
        Fk = std::max(Fk + sc.gapB_tab[j], Mk(i, j - 1) + sc.gapB_tab[j] + alpha);
        max_score = std::max(max_score, Fk);
        const auto &adjlA = bpsA.right_adj[i];
        const auto &adjlB = bpsB.right_adj[j];
        for (auto arcA = adjlA.begin(); (size_t)arcA->left > al; ++arcA)
            for (auto arcB = adjlB.begin(); (size_t)arcB->left > bl; ++arcB) {
                sc_t ns = Mk(arcA->left - 1, arcB->left - 1) + D(*arcA, *arcB);
                if (ns <= max_score) continue;
                max_score = ns;
            }
        return normalize(max_score);
    }

    // aligner.cc:264-368 in state k: the first column and row start at 0 (a
    // free open exclusion) without the global flag of their sequence, at the
    // exclusion score (a closed exclusion of the prefix) with its exclusion
    // flag, else at the indel opening plus gap costs
    void init_state_s(int k, size_t al, size_t ar, size_t bl, size_t br, bool globalA, bool exclA, bool globalB,
                      bool exclB) {
        if (k < 4)
            for (size_t j = bl; j < br; j++) Es(k)[j] = NEG_INF;
        Mat<sc_t> &Mk = Ms(k);
        Mk(al, bl) = 0;
        sc_t indel_score = exclA ? (sc_t)in.p.exclusion : (!globalA ? 0 : alpha);
        size_t i;
        for (i = al + 1; i < ar; i++) {
            if (tc.min_col[i] > bl) break;
            if (!is_neg_inf(indel_score) && !exclA && globalA) indel_score += sc.gapA_tab[i];
            Mk(i, bl) = indel_score;
        }
        for (; i < ar; i++) Mk(i, tc.min_col[i] - 1) = NEG_INF;
        indel_score = exclB ? (sc_t)in.p.exclusion : (!globalB ? 0 : alpha);
        size_t j;
        for (j = bl + 1; j < std::min(br, tc.max_col[al] + 1); j++) {
            if (!is_neg_inf(indel_score) && !exclB && globalB) indel_score += sc.gapB_tab[j];
            Mk(al, j) = indel_score;
        }
        for (i = al + 1; i < ar; i++)
            for (; j < std::min(br, tc.max_col[i] + 1); ++j) Mk(i - 1, j) = NEG_INF;
    }

    // aligner.cc:373-565 with allow_exclusion: all eight states, in the
    // reference's order, each over the whole region
    void align_in_arcmatch_sl(size_t al, size_t ar, size_t bl, size_t br) {
        init_state_s(E_NO_NO, al, ar, bl, br, true, false, true, false);
        init_state_s(E_X_NO, al, ar, bl, br, true, true, true, false);
        init_state_s(E_NO_X, al, ar, bl, br, true, false, true, true);
        init_state_s(E_X_X, al, ar, bl, br, true, true, true, true);
        init_state_s(E_OP_NO, al, ar, bl, br, false, false, true, false);
        init_state_s(E_NO_OP, al, ar, bl, br, true, false, false, false);
        init_state_s(E_X_OP, al, ar, bl, br, true, true, false, false);
        init_state_s(E_OP_X, al, ar, bl, br, false, false, true, true);
        const sc_t ex = in.p.exclusion;
        auto rows = [&](auto cell, bool resetF, int k) {
            for (size_t i = al + 1; i < ar; i++) {
                if (resetF) Fs(k) = NEG_INF;
                const size_t lo = std::max(bl + 1, tc.min_col[i]), hi = std::min(br - 1, tc.max_col[i]);
                for (size_t j = lo; j <= hi; j++) cell(i, j);
            }
        };
        rows([&](size_t i, size_t j) { M(i, j) = align_noex_s(E_NO_NO, al, bl, i, j); }, true, E_NO_NO);
        rows([&](size_t i, size_t j) { Ms(E_OP_NO)(i, j) = std::max(Ms(E_OP_NO)(i - 1, j), M(i, j)); }, false, 0);
        rows([&](size_t i, size_t j) { Ms(E_NO_OP)(i, j) = std::max(Ms(E_NO_OP)(i, j - 1), M(i, j)); }, false, 0);
        rows([&](size_t i, size_t j) {
            Ms(E_NO_X)(i, j) = normalize(std::max(align_noex_s(E_NO_X, al, bl, i, j), Ms(E_NO_OP)(i, j) + ex));
        }, true, E_NO_X);
        rows([&](size_t i, size_t j) { Ms(E_OP_X)(i, j) = std::max(Ms(E_OP_X)(i - 1, j), Ms(E_NO_X)(i, j)); }, false, 0);
        rows([&](size_t i, size_t j) {
            Ms(E_X_NO)(i, j) = normalize(std::max(align_noex_s(E_X_NO, al, bl, i, j), Ms(E_OP_NO)(i, j) + ex));
        }, true, E_X_NO);
        rows([&](size_t i, size_t j) { Ms(E_X_OP)(i, j) = std::max(Ms(E_X_OP)(i, j - 1), Ms(E_X_NO)(i, j)); }, false, 0);
        rows([&](size_t i, size_t j) {
            Ms(E_X_X)(i, j) = normalize(std::max(align_noex_s(E_X_X, al, bl, i, j),
                                                 std::max(Ms(E_OP_X)(i, j) + ex, Ms(E_X_OP)(i, j) + ex)));
        }, true, E_X_X);
    }

    // aligner.cc:574-610
    // with stacking, an arc match whose inner arc match exists and whose
    // arcs are stackable may take the inner one's D plus its stacked score
    void fill_D_entries(int al, int bl) {
        for (int x : ams.left_lists(al, bl)) {
            const ArcMatch &am = ams.ams[x];
            sc_t m = M(am.a.right - 1, am.b.right - 1);
            if (struct_local())  // the best of the non-open states (aligner.cc:586-592)
                for (int k = 1; k < 4; k++) m = std::max(m, Ms(k)(am.a.right - 1, am.b.right - 1));
            D(am) = m + am_score[x];
            if (sc.stacking() && ams.inner[x] >= 0 && am_stackable[x])
                D(am) = std::max(D(am), D(ams.ams[ams.inner[x]]) + am_score_st[x]);
        }
    }

    // aligner.cc:612-657 (no stacking): D of the arc matches with left ends
    // (al-1, bl-1) that have an inner arc match, from the group (al, bl)
    // over the inner region: the inner arc match ends the stack (its region
    // aligned freely) or continues it (its own D)
    void fill_D_entries_noLP(int al, int bl) {
        for (int x : ams.left_lists(al - 1, bl - 1)) {
            const int y = ams.inner[x];
            if (y < 0) continue;
            // with stacking only stackable arc matches, with their stacked
            // score (aligner.cc:624-650)
            if (sc.stacking() && !am_stackable[x]) continue;
            const ArcMatch &am = ams.ams[x], &in_am = ams.ams[y];
            const sc_t m = M(in_am.a.right - 1, in_am.b.right - 1);
            D(am) = std::max(m + am_score[y], D(in_am)) + (sc.stacking() ? am_score_st[x] : am_score[x]);
        }
    }

    // aligner.cc:659-732
    void align_D() {
        const bool noLP = in.p.no_lonely_pairs != 0;
        for (size_t al = lenA + 1; al > 1;) {
            al--;
            size_t max_bl = std::min((size_t)lenB, tc.max_col[al]);
            size_t min_bl = std::max((size_t)1, tc.min_col[al]);
            for (size_t bl = max_bl + 1; bl > min_bl;) {
                bl--;
                if (!tc.is_valid_match(al, bl)) continue;
                int max_ar = (int)al, max_br = (int)bl;
                if (noLP)
                    ams.max_right_ends_noLP((int)al, (int)bl, &max_ar, &max_br);
                else
                    ams.max_right_ends((int)al, (int)bl, &max_ar, &max_br);
                if ((int)al == max_ar || (int)bl == max_br) continue;
                if (counting) cnt_G++;
                if (struct_local())
                    align_in_arcmatch_sl(al, max_ar, bl, max_br);
                else
                    align_in_arcmatch(al, max_ar, bl, max_br);
                if (noLP)
                    fill_D_entries_noLP((int)al, (int)bl);
                else
                    fill_D_entries((int)al, (int)bl);
            }
        }
    }

    // aligner.cc:736-819: column init free with the second sequence's left
    // flag, row init with the first's; with a right flag the best of the
    // last column (second sequence's right end, rows ascending) and then of
    // the last row (first sequence's right end, columns ascending), first
    // strict maximum (no anchors: the rightmost anchor is (0, 0))
    sc_t align_top_level_global() {
        const int fe = in.p.free_endgaps;
        const size_t sA = rsA, eA = reA, sB = rsB, eB = reB;
        init_state(sA - 1, eA + 1, sB - 1, eB + 1, !fe_left_2(fe), !fe_left_1(fe), lam);
        for (size_t i = sA; i <= eA; i++) {
            F = NEG_INF;
            size_t min_col = std::max(sB, tc.min_col[i]);
            size_t max_col = std::min(eB, tc.max_col[i]);
            for (size_t j = min_col; j <= max_col; j++) M(i, j) = align_noex(sA - 1, sB - 1, i, j, lam);
        }
        // the end cell of a proper restriction outside the band: minus
        // infinity (the reference reads whatever an earlier computation left
        // in its matrix there); the whole sequences keep the matrix's value
        const bool whole = sA == 1 && sB == 1 && eA == (size_t)lenA && eB == (size_t)lenB;
        sc_t max_score = (whole || tc.is_valid(eA, eB)) ? M(eA, eB) : NEG_INF;
        max_i = (int)eA;
        max_j = (int)eB;
        if (fe_right_1(fe) || fe_right_2(fe)) {
            max_score = NEG_INF;
            max_i = (int)sA - 1;
            max_j = (int)sB - 1;
        }
        if (fe_right_2(fe)) {
            for (size_t i = sA; i <= eA; i++) {
                if (tc.max_col[i] >= eB && M(i, eB) > max_score) {
                    max_score = M(i, eB);
                    max_i = (int)i;
                    max_j = (int)eB;
                }
            }
        }
        if (fe_right_1(fe)) {
            size_t min_col = std::max(sB, tc.min_col[eA]);
            size_t max_col = std::min(eB, tc.max_col[eA]);
            for (size_t j = min_col; j <= max_col; j++) {
                if (M(eA, j) > max_score) {
                    max_score = M(eA, j);
                    max_i = (int)eA;
                    max_j = (int)j;
                }
            }
        }
        return max_score;
    }

    // aligner.cc:823-880 (no anchors: left anchor (lenA+1,lenB+1), right (0,0))
    sc_t align_top_level_locally() {
        const size_t sA = rsA, eA = reA, sB = rsB, eB = reB;
        sc_t max_score = 0;
        max_i = (int)sA - 1;
        max_j = (int)sB - 1;
        init_state(sA - 1, eA + 1, sB - 1, eB + 1, false, false, lam);
        for (size_t i = sA; i <= eA; i++) {
            F = NEG_INF;
            size_t min_col = std::max(sB, tc.min_col[i]);
            size_t max_col = std::min(eB, tc.max_col[i]);
            for (size_t j = min_col; j <= max_col; j++) {
                M(i, j) = align_noex(sA - 1, sB - 1, i, j, lam);
                M(i, j) = std::max((sc_t)0, M(i, j));
                if (max_score < M(i, j)) {
                    max_score = M(i, j);
                    max_i = (int)i;
                    max_j = (int)j;
                }
            }
        }
        return max_score;
    }

    sc_t align() {
        align_D();
        D_done = true;
        return in.p.sequ_local ? align_top_level_locally() : align_top_level_global();
    }

    // aligner.cc:1603-1622: the top level with position penalty P (local, or
    // global with free end gaps), then the traceback
    sc_t penalized_align(sc_t P) {
        if (!D_done) align_D();
        D_done = true;
        lam = P;
        sc_t score = in.p.sequ_local ? align_top_level_locally() : align_top_level_global();
        trace();
        return score;
    }

    // aligner.cc:1526-1593: Dinkelbach iterations over the local top level
    // with penalty lambda; length = max_i - min_i + 1 + max_j - min_j + 1
    // of the traced alignment; lambda <- (score + length lambda) /
    // (length + L) (integer division) until it repeats
    sc_t normalized_align(sc_t L) {
        if (!in.p.sequ_local) throw std::runtime_error("normalized alignment needs sequence-local alignment");
        if (!D_done) align_D();
        D_done = true;
        sc_t new_lambda = 0, lambda = -1;
        while (lambda != new_lambda) {
            lambda = new_lambda;
            lam = lambda;
            sc_t score = align_top_level_locally();
            trace();
            const sc_t length = (max_i - min_i + 1) + (max_j - min_j + 1);
            score += length * lambda;
            new_lambda = score / (length + L);
        }
        return new_lambda;
    }

    // ---- traceback: aligner.cc:967-1028, 1082-1229, 1233-1343, 1345-1358
    void append(int a, int b) { edges.push_back({a, b}); }

    // aligner.cc:967-1028: with stacking, an arc match whose D is the inner
    // one's plus its stacked score traces the inner arc match directly (the
    // reference compares with the stacked score of any arc match with an
    // inner one; for arcs that are not stackable that score is undefined
    // there and never matches, so only stackable ones are compared here)
    void trace_arcmatch(const ArcMatch &am) {
        const int y = ams.inner[am.idx];
        if (sc.stacking() && y >= 0 && am_stackable[am.idx] && D(am) == D(ams.ams[y]) + am_score_st[am.idx]) {
            const ArcMatch &in_am = ams.ams[y];
            structA[in_am.a.left - 1] = '(';
            structA[in_am.a.right - 1] = ')';
            structB[in_am.b.left - 1] = '(';
            structB[in_am.b.right - 1] = ')';
            append(in_am.a.left, in_am.b.left);
            trace_arcmatch(in_am);
            append(in_am.a.right, in_am.b.right);
            return;
        }
        if (struct_local()) {
            // aligner.cc:1012-1025: the first non-open state whose value
            // gives D(am)
            align_in_arcmatch_sl(am.a.left, am.a.right, am.b.left, am.b.right);
            for (int k = 0; k < 4; k++)
                if (D(am) == Ms(k)(am.a.right - 1, am.b.right - 1) + am_score[am.idx]) {
                    trace_in_arcmatch_s(k, am.a.left, am.a.right - 1, am.b.left, am.b.right - 1);
                    break;
                }
            return;
        }
        align_in_arcmatch(am.a.left, am.a.right, am.b.left, am.b.right);
        trace_in_arcmatch(am.a.left, am.a.right - 1, am.b.left, am.b.right - 1, false);
    }

    // aligner.cc:1233-1343 inside an arc match in state k (--struct-local):
    // excluded positions get no edge; a prefix run out in a sequence with an
    // exclusion (open or closed) is excluded rather than padded with gaps
    void trace_in_arcmatch_s(int k, size_t al, size_t i, size_t bl, size_t j) {
        Mat<sc_t> &Mk = Ms(k);
        const sc_t ex = in.p.exclusion;
        if (i <= al) {
            if (k == E_NO_NO || k == E_OP_NO || k == E_X_NO)
                for (size_t t = bl + 1; t <= j; t++) append(-1, (int)t);
            return;
        }
        if (j <= bl) {
            if (k == E_NO_NO || k == E_NO_OP || k == E_NO_X)
                for (size_t t = al + 1; t <= i; t++) append((int)t, -1);
            return;
        }
        switch (k) {
            case E_NO_NO: trace_noex_s(k, al, i, bl, j); break;
            case E_OP_NO:
                if (Mk(i, j) == Mk(i - 1, j)) trace_in_arcmatch_s(k, al, i - 1, bl, j);
                else if (Mk(i, j) == M(i, j)) trace_in_arcmatch_s(E_NO_NO, al, i, bl, j);
                break;
            case E_NO_OP:
                if (Mk(i, j) == Mk(i, j - 1)) trace_in_arcmatch_s(k, al, i, bl, j - 1);
                else if (Mk(i, j) == M(i, j)) trace_in_arcmatch_s(E_NO_NO, al, i, bl, j);
                break;
            case E_NO_X:
                if (Mk(i, j) == Ms(E_NO_OP)(i, j) + ex) trace_in_arcmatch_s(E_NO_OP, al, i, bl, j);
                else trace_noex_s(k, al, i, bl, j);
                break;
            case E_OP_X:
                if (Mk(i, j) == Mk(i - 1, j)) trace_in_arcmatch_s(k, al, i - 1, bl, j);
                else if (Mk(i, j) == Ms(E_NO_X)(i, j)) trace_in_arcmatch_s(E_NO_X, al, i, bl, j);
                break;
            case E_X_NO:
                if (Mk(i, j) == Ms(E_OP_NO)(i, j) + ex) trace_in_arcmatch_s(E_OP_NO, al, i, bl, j);
                else trace_noex_s(k, al, i, bl, j);
                break;
            case E_X_OP:
                if (Mk(i, j) == Mk(i, j - 1)) trace_in_arcmatch_s(k, al, i, bl, j - 1);
                else if (Mk(i, j) == Ms(E_X_NO)(i, j)) trace_in_arcmatch_s(E_X_NO, al, i, bl, j);
                break;
            case E_X_X:
                if (Mk(i, j) == Ms(E_OP_X)(i, j) + ex) trace_in_arcmatch_s(E_OP_X, al, i, bl, j);
                else if (Mk(i, j) == Ms(E_X_OP)(i, j) + ex) trace_in_arcmatch_s(E_X_OP, al, i, bl, j);
                else trace_noex_s(k, al, i, bl, j);
                break;
        }
    }

    // aligner.cc:1082-1229 in state k < 4 inside an arc match (no position
    // penalty, not the top level)
    void trace_noex_s(int k, size_t oal, size_t i, size_t obl, size_t j) {
        Mat<sc_t> &Mk = Ms(k);
        if (tc.is_valid(i - 1, j - 1) && Mk(i, j) == Mk(i - 1, j - 1) + sc.sigma_tab(i, j)) {
            trace_in_arcmatch_s(k, oal, i - 1, obl, j - 1);
            append((int)i, (int)j);
            return;
        }
        if (alpha == 0) {
            if (tc.is_valid(i - 1, j) && Mk(i, j) == Mk(i - 1, j) + sc.gapA_tab[i]) {
                trace_in_arcmatch_s(k, oal, i - 1, obl, j);
                append((int)i, -1);
                return;
            }
            if (tc.is_valid(i, j - 1) && Mk(i, j) == Mk(i, j - 1) + sc.gapB_tab[j]) {
                trace_in_arcmatch_s(k, oal, i, obl, j - 1);
                append(-1, (int)j);
                return;
            }
        } else {
            sc_t gap_cost = alpha;
            for (size_t t = 1; i >= oal + t; t++) {
                if (!tc.is_valid(i - t, j)) break;
                gap_cost += sc.gapA_tab[i - t + 1];
                if (Mk(i, j) == Mk(i - t, j) + gap_cost) {
                    trace_in_arcmatch_s(k, oal, i - t, obl, j);
                    for (size_t l = t; l > 0; l--) append((int)(i - l + 1), -1);
                    return;
                }
            }
            gap_cost = alpha;
            for (size_t t = 1; j >= obl + t; t++) {
                if (!tc.is_valid(i, j - t)) break;
                gap_cost += sc.gapB_tab[j - t + 1];
                if (Mk(i, j) == Mk(i, j - t) + gap_cost) {
                    trace_in_arcmatch_s(k, oal, i, obl, j - t);
                    for (size_t l = t; l > 0; l--) append(-1, (int)(j - l + 1));
                    return;
                }
            }
        }
        if (!tc.is_valid(i - 1, j - 1)) return;
        for (int x : ams.right_lists(i, j)) {
            const ArcMatch &am = ams.ams[x];
            if ((size_t)am.a.left <= oal || (size_t)am.b.left <= obl) continue;
            size_t al = am.a.left, bl = am.b.left;
            if (Mk(i, j) == Mk(al - 1, bl - 1) + D(am)) {
                trace_in_arcmatch_s(k, oal, al - 1, obl, bl - 1);
                structA[al - 1] = '(';
                structA[i - 1] = ')';
                structB[bl - 1] = '(';
                structB[j - 1] = ')';
                append((int)al, (int)bl);
                trace_arcmatch(am);
                append((int)i, (int)j);
                return;
            }
        }
    }

    // aligner.cc:1030-1078 (no stacking, not structure-local): the inner
    // arc match of the stack, which continues the stack when D(am) ==
    // D(inner) + s(am), else ends it with its region traced
    void trace_arcmatch_noLP(const ArcMatch &am) {
        const ArcMatch &y = ams.ams[ams.inner[am.idx]];
        structA[y.a.left - 1] = '(';
        structA[y.a.right - 1] = ')';
        structB[y.b.left - 1] = '(';
        structB[y.b.right - 1] = ')';
        append(y.a.left, y.b.left);
        if (D(am) == D(y) + (sc.stacking() ? am_score_st[am.idx] : am_score[am.idx])) {
            trace_arcmatch_noLP(y);
        } else {
            align_in_arcmatch(y.a.left, y.a.right, y.b.left, y.b.right);
            trace_in_arcmatch(y.a.left, y.a.right - 1, y.b.left, y.b.right - 1, false);
        }
        append(y.a.right, y.b.right);
    }

    void trace_noex(size_t oal, size_t i, size_t obl, size_t j, bool tl) {
        const sc_t lm = tl ? lam : 0;
        if (tc.is_valid(i - 1, j - 1) && M(i, j) == M(i - 1, j - 1) + (sc.sigma_tab(i, j) - 2 * lm)) {
            trace_in_arcmatch(oal, i - 1, obl, j - 1, tl);
            append((int)i, (int)j);
            return;
        }
        if (alpha == 0) {
            if (tc.is_valid(i - 1, j) && M(i, j) == M(i - 1, j) + (sc.gapA_tab[i] - lm)) {
                trace_in_arcmatch(oal, i - 1, obl, j, tl);
                append((int)i, -1);
                return;
            }
            if (tc.is_valid(i, j - 1) && M(i, j) == M(i, j - 1) + (sc.gapB_tab[j] - lm)) {
                trace_in_arcmatch(oal, i, obl, j - 1, tl);
                append(-1, (int)j);
                return;
            }
        } else {
            sc_t gap_cost = alpha;
            for (size_t k = 1; i >= oal + k; k++) {
                if (!tc.is_valid(i - k, j)) break;
                gap_cost += sc.gapA_tab[i - k + 1] - lm;
                if (M(i, j) == M(i - k, j) + gap_cost) {
                    trace_in_arcmatch(oal, i - k, obl, j, tl);
                    for (size_t l = k; l > 0; l--) append((int)(i - l + 1), -1);
                    return;
                }
            }
            gap_cost = alpha;
            for (size_t k = 1; j >= obl + k; k++) {
                if (!tc.is_valid(i, j - k)) break;
                gap_cost += sc.gapB_tab[j - k + 1] - lm;
                if (M(i, j) == M(i, j - k) + gap_cost) {
                    trace_in_arcmatch(oal, i, obl, j - k, tl);
                    for (size_t l = k; l > 0; l--) append(-1, (int)(j - l + 1));
                    return;
                }
            }
        }
        if (!tc.is_valid(i - 1, j - 1)) return;
        for (int x : ams.right_lists(i, j)) {
            const ArcMatch &am = ams.ams[x];
            if ((size_t)am.a.left <= oal || (size_t)am.b.left <= obl) continue;
            size_t al = am.a.left, bl = am.b.left;
            if (M(i, j) == M(al - 1, bl - 1) + D(am) -
                               lm * (sc_t)((am.a.right - am.a.left + 1) + (am.b.right - am.b.left + 1))) {
                trace_in_arcmatch(oal, al - 1, obl, bl - 1, tl);
                structA[al - 1] = '(';
                structA[i - 1] = ')';
                structB[bl - 1] = '(';
                structB[j - 1] = ')';
                append((int)al, (int)bl);
                if (in.p.no_lonely_pairs)
                    trace_arcmatch_noLP(am);
                else
                    trace_arcmatch(am);
                append((int)i, (int)j);
                return;
            }
        }
    }

    void trace_in_arcmatch(size_t al, size_t i, size_t bl, size_t j, bool tl) {
        if (tl && in.p.sequ_local && M(i, j) == 0) {
            min_i = (int)i;
            min_j = (int)j;
            return;
        }
        // aligner.cc:1260-1289: no padding of free leading end gaps
        if (i <= al) {
            if (!(tl && (in.p.sequ_local || fe_left_1(in.p.free_endgaps))))
                for (size_t k = bl + 1; k <= j; k++) append(-1, (int)k);
            return;
        }
        if (j <= bl) {
            if (!(tl && (in.p.sequ_local || fe_left_2(in.p.free_endgaps))))
                for (size_t k = al + 1; k <= i; k++) append((int)k, -1);
            return;
        }
        trace_noex(al, i, bl, j, tl);
    }

    void trace() {  // alignment_.clear() first (aligner.cc:1345-1358)
        edges.clear();
        structA.assign(lenA, '.');
        structB.assign(lenB, '.');
        trace_in_arcmatch(rsA - 1, max_i, rsB - 1, max_j, true);
    }

    // Aligner::suboptimal (aligner.cc:1382-1514), not normalized: the best
    // alignment of a restriction, then the restriction split at the middle
    // of the traced alignment's extent in the longer sequence (both halves
    // keep the middle position), both halves queued with their scores; a
    // max-heap on the score (greater_second, aligner.hh:34-45, compares
    // with <), k alignments or until the best queued score is at most the
    // threshold.  Empty alignments are skipped.
    struct Subopt {
        sc_t score;
        std::vector<std::pair<int, int>> edges;
        std::string structA, structB;
        int sA, sB, eA, eB;  // its restriction
    };
    std::vector<Subopt> suboptimal(int k, sc_t threshold) {
        struct Restr {
            int sA, sB, eA, eB;
        };
        typedef std::pair<Restr, sc_t> task_t;
        struct Less {
            bool operator()(const task_t &a, const task_t &b) const { return a.second < b.second; }
        };
        std::vector<Subopt> out;
        if (!D_done) align_D();
        D_done = true;
        auto top = [&](const Restr &r) {
            rsA = r.sA;
            rsB = r.sB;
            reA = r.eA;
            reB = r.eB;
            return in.p.sequ_local ? align_top_level_locally() : align_top_level_global();
        };
        std::priority_queue<task_t, std::vector<task_t>, Less> tasks;
        const Restr whole{1, 1, lenA, lenB};
        tasks.push(task_t(whole, top(whole)));
        size_t i = 1;
        while (k < 0 || i <= (size_t)k) {
            if (tasks.empty()) break;
            task_t task = tasks.top();
            tasks.pop();
            const Restr r = task.first;
            const sc_t task_score = task.second;
            // a restriction without an alignment (minus infinity) ends the
            // enumeration like one below the threshold
            if (task_score < threshold + 1 || is_neg_inf(task_score)) break;
            top(r);
            trace();
            if (edges.empty()) continue;
            out.push_back({task_score, edges, structA, structB, r.sA, r.sB, r.eA, r.eB});
            if (k >= 0 && i == (size_t)k) break;
            // start / end positions (alignment.cc:170-195)
            int fa = -1, fb = -1, la = 0, lb = 0;
            for (auto &e : edges) {
                if (fa < 0 && e.first > 0) fa = e.first;
                if (fb < 0 && e.second > 0) fb = e.second;
                if (e.first > 0) la = e.first;
                if (e.second > 0) lb = e.second;
            }
            if (fa < 0) fa = lenA;
            if (fb < 0) fb = lenB;
            Restr r1 = r, r2 = r;
            if (r.eA - r.sA > r.eB - r.sB) {
                const int split = (fa + la) / 2;
                r1.eA = split;
                r2.sA = split;
            } else {
                const int split = (fb + lb) / 2;
                r1.eB = split;
                r2.sB = split;
            }
            // a restriction left without rows or columns aligns nothing
            auto score_of = [&](const Restr &x) -> sc_t {
                if (x.eA < x.sA || x.eB < x.sB || x.sA < 1 || x.sB < 1 || x.eA > lenA || x.eB > lenB)
                    return in.p.sequ_local ? 0 : NEG_INF;
                return top(x);
            };
            const sc_t s1 = score_of(r1);
            const sc_t s2 = score_of(r2);
            tasks.push(task_t(r1, s1));
            tasks.push(task_t(r2, s2));
            ++i;
        }
        rsA = 1;
        rsB = 1;
        reA = lenA;
        reB = lenB;
        return out;
    }
};


// ---------------------------------------------------------------- AlignerP
// Inside: aligner_p.icc:148-438 (init_M, init_E, comp_E/F/M_entry,
// align_inside_arcmatch, fill_D, align_D, align_inside), pf_score_t = double.
// Outside quantities: reverse-mode adjoint of the inside recursion, group by
// group in reverse order; Dbar(am) = dZ/dD(am) gives P(am) = D Dbar / Z, the
// quantity aligner_p.icc:1150-1195 computes as D D' pf_scale / (Z e^{s/T}).
// Base-match probabilities follow aligner_p.icc:1201-1399: match edges
// directly inside arc matches with P(am) > sqrt(min_am_prob), plus the top
// level, optionally plus P(am) at both arc ends.
struct AlignerP {
    Inputs &in;
    const Band &tc;
    const ArcMatches &ams;
    const Scoring &sc;
    const BasePairs &bpsA, &bpsB;
    int lenA, lenB;
    double T, ps;
    Mat<double> esig;
    std::vector<double> egA, egB, eam;
    double eo;
    Mat<double> M;
    std::vector<double> E;
    double F = 0;
    Mat<double> Dmat;  // N_A x N_B, 0 where invalid (aligner_p.icc:25)
    double Z = 0;

    explicit AlignerP(Inputs &in_)
        : in(in_), tc(*in_.band), ams(*in_.am), sc(*in_.sc), bpsA(*in_.bpsA), bpsB(*in_.bpsB),
          lenA(in_.A.length()), lenB(in_.B.length()), T((double)in_.p.temperature_alipf), ps(in_.p.pf_scale) {
        // scoring.hh:882-931: boltzmann_weight(s) = exp(s / (pf_score_t)T)
        esig.resize(lenA + 1, lenB + 1, 0.0);
        for (int i = 1; i <= lenA; ++i)
            for (int j = 1; j <= lenB; ++j) esig(i, j) = exp(sc.sigma_tab(i, j) / T);
        egA.assign(lenA + 1, 0.0);
        egB.assign(lenB + 1, 0.0);
        for (int i = 1; i <= lenA; ++i) egA[i] = exp(sc.gapA_tab[i] / T);
        for (int j = 1; j <= lenB; ++j) egB[j] = exp(sc.gapB_tab[j] / T);
        eo = exp(in.p.indel_opening / T);
        eam.resize(ams.ams.size());
        for (size_t k = 0; k < ams.ams.size(); ++k) eam[k] = exp(sc.arcmatch(ams.ams[k]) / T);
        M.resize(lenA + 1, lenB + 1, 0.0);
        E.assign(lenB + 1, 0.0);
        Dmat.resize(bpsA.arcs.size(), bpsB.arcs.size(), 0.0);
    }
    double D(const Arc &a, const Arc &b) const { return Dmat(a.idx, b.idx); }

    void init_M(size_t al, size_t ar, size_t bl, size_t br) {  // aligner_p.icc:150-189
        M(al, bl) = 1.0 / ps;
        double indel_score = eo / ps;
        size_t i;
        for (i = al + 1; i < ar; i++) {
            if (tc.min_col[i] > bl) break;
            indel_score *= egA[i];
            M(i, bl) = indel_score;
        }
        for (; i < ar; i++) M(i, tc.min_col[i] - 1) = 0;
        size_t max_col = std::min(br - 1, tc.max_col[al]);
        indel_score = eo / ps;
        size_t j;
        for (j = bl + 1; j <= max_col; j++) {
            indel_score *= egB[j];
            M(al, j) = indel_score;
        }
        for (size_t i2 = al + 1; i2 < ar; i2++)
            for (; j < std::min(br, tc.max_col[i2] + 1); ++j) M(i2 - 1, j) = 0;
    }
    double comp_M_entry(size_t al, size_t bl, size_t i, size_t j) {  // aligner_p.icc:234-278
        double pf = M(i - 1, j - 1) * esig(i, j) + E[j] + F;
        const auto &adjlA = bpsA.right_adj[i];
        const auto &adjlB = bpsB.right_adj[j];
        for (auto a = adjlA.begin(); (size_t)a->left > al; ++a)
            for (auto b = adjlB.begin(); (size_t)b->left > bl; ++b)
                pf += M(a->left - 1, b->left - 1) * D(*a, *b) * ps;
        return pf;
    }
    void align_inside_arcmatch(size_t al, size_t ar, size_t bl, size_t br) {  // aligner_p.icc:285-312
        init_M(al, ar, bl, br);
        for (size_t j = bl; j < br; j++) E[j] = 0;
        for (size_t i = al + 1; i < ar; i++) {
            F = 0;
            size_t min_col = std::max(bl + 1, tc.min_col[i]);
            size_t max_col = std::min(br - 1, tc.max_col[i]);
            for (size_t j = min_col; j <= max_col; j++) {
                E[j] = E[j] * egA[i] + (M(i - 1, j) - E[j]) * egA[i] * eo;
                F = F * egB[j] + (M(i, j - 1) - F) * egB[j] * eo;
                M(i, j) = comp_M_entry(al, bl, i, j);
            }
        }
    }
    void max_right_ends(int al, int bl, int *ar, int *br) const { ams.max_right_ends(al, bl, ar, br); }
    double inside() {  // aligner_p.icc:358-438
        for (size_t al = lenA; al >= 1; al--) {
            size_t min_bl = std::max((size_t)1, tc.min_col[al]);
            size_t max_bl = std::min((size_t)lenB, tc.max_col[al]);
            for (size_t bl = max_bl; bl >= min_bl; bl--) {
                int max_ar = (int)al, max_br = (int)bl;
                max_right_ends((int)al, (int)bl, &max_ar, &max_br);
                align_inside_arcmatch(al, max_ar, bl, max_br);
                for (int x : ams.left_lists(al, bl)) {
                    const ArcMatch &am = ams.ams[x];
                    Dmat(am.a.idx, am.b.idx) = M(am.a.right - 1, am.b.right - 1) * eam[x];
                }
            }
        }
        align_inside_arcmatch(0, lenA + 1, 0, lenB + 1);
        Z = M(lenA, lenB);
        return Z;
    }

    // ---- adjoint of align_inside_arcmatch for one group.  Recomputes the
    // group's forward values, then sweeps cells in reverse order.  Seeds:
    // Mbar at cells.  Accumulates into Dbar (if given), match-edge weights
    // (edge(i,j) += M(i-1,j-1) esig Mbar(i,j)) and gap-opening weight.
    struct Fwd {
        size_t al, ar, bl, br;
        Mat<double> Mv, Ev, Fv;  // values after the cell update
    };
    void forward_group(size_t al, size_t ar, size_t bl, size_t br, Fwd &f) {
        f.al = al; f.ar = ar; f.bl = bl; f.br = br;
        align_inside_arcmatch(al, ar, bl, br);
        // re-run capturing E, F per cell (values identical to the fill above)
        f.Mv.resize(lenA + 1, lenB + 1, 0.0);
        f.Ev.resize(lenA + 1, lenB + 1, 0.0);
        f.Fv.resize(lenA + 1, lenB + 1, 0.0);
        for (size_t i = al; i < ar; ++i)
            for (size_t j = bl; j < br; ++j) f.Mv(i, j) = M(i, j);
        std::vector<double> e(lenB + 1, 0.0);
        for (size_t i = al + 1; i < ar; i++) {
            double fv = 0;
            size_t min_col = std::max(bl + 1, tc.min_col[i]);
            size_t max_col = std::min(br - 1, tc.max_col[i]);
            for (size_t j = min_col; j <= max_col; j++) {
                e[j] = e[j] * egA[i] + (M(i - 1, j) - e[j]) * egA[i] * eo;
                fv = fv * egB[j] + (M(i, j - 1) - fv) * egB[j] * eo;
                f.Ev(i, j) = e[j];
                f.Fv(i, j) = fv;
            }
        }
    }
    // fragment capture of reverse_group: at the visit of cell (j, l), adds
    // M(i-1, k-1) of the group times the cell's adjoint to *out
    struct FragCap {
        size_t i, j, k, l;
        double *out;
    };
    // cell (x, y) holds a value of the group's fill: its init row and column
    // (init_M) or a computed cell
    bool live(const Fwd &f, size_t x, size_t y) const {
        if (x < f.al || x >= f.ar || y < f.bl || y >= f.br) return false;
        if (x == f.al) return y <= std::min(f.br - 1, tc.max_col[f.al]);
        if (y == f.bl) return tc.min_col[x] <= f.bl;
        return y >= tc.min_col[x] && y <= tc.max_col[x];
    }
    // Mbar: seeds on entry (indexed like M), consumed
    void reverse_group(const Fwd &f, Mat<double> &Mbar, std::vector<double> *Dbar, Mat<double> *edge,
                       double *openings, const FragCap *cap = nullptr) {
        const size_t al = f.al, ar = f.ar, bl = f.bl, br = f.br;
        std::vector<double> Ebar(lenB + 1, 0.0);
        for (size_t i = ar - 1; i > al; --i) {
            double Fbar = 0;
            size_t min_col = std::max(bl + 1, tc.min_col[i]);
            size_t max_col = std::min(br - 1, tc.max_col[i]);
            if (min_col > max_col) {
                for (size_t j = bl; j < br; ++j) Ebar[j] = 0;  // E of empty row: chain broken? keep semantics
                continue;
            }
            for (size_t j = max_col; j >= min_col; --j) {
                const double g = Mbar(i, j);
                if (cap && i == cap->j && j == cap->l && live(f, cap->i - 1, cap->k - 1))
                    *cap->out += g * f.Mv(cap->i - 1, cap->k - 1);
                // M(i,j) = M(i-1,j-1) esig + E + F + arcs
                Mbar(i - 1, j - 1) += g * esig(i, j);
                if (edge) (*edge)(i, j) += g * f.Mv(i - 1, j - 1) * esig(i, j);
                const double Eb = Ebar[j] + g;
                const double Fb = Fbar + g;
                const auto &adjlA = bpsA.right_adj[i];
                const auto &adjlB = bpsB.right_adj[j];
                for (auto a = adjlA.begin(); (size_t)a->left > al; ++a)
                    for (auto b = adjlB.begin(); (size_t)b->left > bl; ++b) {
                        const double dv = D(*a, *b);
                        if (dv == 0) continue;
                        Mbar(a->left - 1, b->left - 1) += g * dv * ps;
                        if (Dbar) {
                            int k = am_index(a->idx, b->idx);
                            if (k >= 0) (*Dbar)[k] += g * f.Mv(a->left - 1, b->left - 1) * ps;
                        }
                    }
                // F(i,j) = Fprev gB (1-o) + M(i,j-1) gB o
                const double Mleft = f.Mv(i, j - 1);
                const double Fprev = (j > min_col) ? f.Fv(i, j - 1) : 0.0;
                Mbar(i, j - 1) += Fb * egB[j] * eo;
                if (openings) *openings += Fb * (Mleft - Fprev) * egB[j] * eo;
                Fbar = Fb * egB[j] * (1 - eo);
                // E(i,j) = Eprev gA (1-o) + M(i-1,j) gA o
                const double Eprev = (i - 1 > al && f.Ev(i - 1, j) != 0) ? f.Ev(i - 1, j) : eprev_at(f, i, j);
                const double Mup = f.Mv(i - 1, j);
                Mbar(i - 1, j) += Eb * egA[i] * eo;
                if (openings) *openings += Eb * (Mup - Eprev) * egA[i] * eo;
                Ebar[j] = Eb * egA[i] * (1 - eo);
                if (j == min_col) break;
            }
            // columns of row i not computed: their E chain ends here
            for (size_t j = bl; j < br; ++j)
                if (j < min_col || j > max_col) Ebar[j] = 0;
        }
        // init row al and column bl carry one opening each (init_M)
        if (openings) {
            for (size_t j = bl + 1; j < br && j <= tc.max_col[al]; ++j) *openings += Mbar(al, j) * f.Mv(al, j);
            for (size_t i = al + 1; i < ar && tc.min_col[i] <= bl; ++i) *openings += Mbar(i, bl) * f.Mv(i, bl);
        }
    }
    // E value entering row i at column j (E of the last row above that
    // computed column j, 0 if none)
    double eprev_at(const Fwd &f, size_t i, size_t j) const {
        return (i - 1 > f.al) ? f.Ev(i - 1, j) : 0.0;
    }
    std::vector<int> am_of;  // arcA idx * N_B + arcB idx -> am index
    int am_index(int a, int b) {
        if (am_of.empty()) {
            am_of.assign(bpsA.arcs.size() * bpsB.arcs.size(), -1);
            for (size_t k = 0; k < ams.ams.size(); ++k) am_of[ams.ams[k].a.idx * bpsB.arcs.size() + ams.ams[k].b.idx] = (int)k;
        }
        return am_of[(size_t)a * bpsB.arcs.size() + b];
    }

    // Fragment match probability, AlignerP::compute_fragment_match_prob
    // (aligner_p.icc:1440-1479): the weight of the structure alignments that
    // align A[i..j] with B[k..l] as one block, over Z.  The block's own
    // alignments are the inside of the virtual arc match (i-1, j+1; k-1,
    // l+1): in = M(j, l) of that region.  The rest: the alignment passes
    // through the cells (i-1, k-1) and (j, l) of one group (an arc match's
    // interior or the top level), with no arc match crossing either; its
    // weight is out = sum over the groups R holding both cells of
    // M_R(i-1, k-1) times the adjoint of M_R(j, l) at the visit of the cell
    // in the reverse sweep (the weight of every completion from (j, l) to
    // the group's seeds), which the reference computes as the outside of
    // the hole [i..j] x [k..l] (Mprime, virtual_Mprime; aligner_p.icc:
    // 786-1046).  P = in out pf_scale / Z.  The block starts and ends its
    // own gap runs: a run through either cell counts as two, as in the
    // reference's decomposition.  Dbar: the outside pass's adjoints.
    double fragment(const std::vector<double> &Dbar, size_t i, size_t j, size_t k, size_t l) {
        if (i < 1 || k < 1 || i > j || k > l || j > (size_t)lenA || l > (size_t)lenB)
            throw std::runtime_error("fragment positions out of range");
        double fin = 0.0;
        {
            align_inside_arcmatch(i - 1, j + 1, k - 1, l + 1);
            Fwd v;
            v.al = i - 1; v.ar = j + 1; v.bl = k - 1; v.br = l + 1;
            if (j > i - 1 && l > k - 1 && live(v, j, l)) fin = M(j, l);
        }
        double out = 0.0;
        const FragCap cap{i, j, k, l, &out};
        Fwd f;
        Mat<double> Mbar;
        forward_group(0, lenA + 1, 0, lenB + 1, f);
        Mbar.resize(lenA + 1, lenB + 1, 0.0);
        Mbar(lenA, lenB) = 1.0;
        reverse_group(f, Mbar, nullptr, nullptr, nullptr, &cap);
        for (size_t al = 1; al + 1 <= i; ++al) {
            size_t min_bl = std::max((size_t)1, tc.min_col[al]);
            size_t max_bl = std::min((size_t)lenB, tc.max_col[al]);
            for (size_t bl = min_bl; bl <= max_bl && bl + 1 <= k; ++bl) {
                const auto &lst = ams.left_lists(al, bl);
                if (lst.empty()) continue;
                int max_ar = (int)al, max_br = (int)bl;
                max_right_ends((int)al, (int)bl, &max_ar, &max_br);
                if ((size_t)max_ar <= j || (size_t)max_br <= l) continue;
                forward_group(al, max_ar, bl, max_br, f);
                Mbar.resize(lenA + 1, lenB + 1, 0.0);
                for (int x : lst) {
                    const ArcMatch &am = ams.ams[x];
                    Mbar(am.a.right - 1, am.b.right - 1) += Dbar[x] * eam[x];
                }
                reverse_group(f, Mbar, nullptr, nullptr, nullptr, &cap);
            }
        }
        return fin * out * ps / Z;
    }

    // Outside pass; returns Dbar (per am), full match-edge weights, restricted
    // match-edge weights (reference base-match semantics), openings weight.
    void outside(std::vector<double> &Dbar, Mat<double> &edge_full, Mat<double> &edge_ref, double &openings) {
        const size_t nA = ams.ams.size();
        Dbar.assign(nA, 0.0);
        edge_full.resize(lenA + 1, lenB + 1, 0.0);
        edge_ref.resize(lenA + 1, lenB + 1, 0.0);
        openings = 0;
        Mat<double> Mbar;
        Fwd f;
        // top level: Z = M(n, m)
        forward_group(0, lenA + 1, 0, lenB + 1, f);
        Mbar.resize(lenA + 1, lenB + 1, 0.0);
        Mbar(lenA, lenB) = 1.0;
        Mat<double> Mbar2 = Mbar;
        reverse_group(f, Mbar, &Dbar, &edge_full, &openings);
        reverse_group(f, Mbar2, nullptr, &edge_ref, nullptr);
        const double thr = sqrt(in.p.min_am_prob);
        for (size_t al = 1; al <= (size_t)lenA; ++al) {
            size_t min_bl = std::max((size_t)1, tc.min_col[al]);
            size_t max_bl = std::min((size_t)lenB, tc.max_col[al]);
            for (size_t bl = min_bl; bl <= max_bl; ++bl) {
                const auto &lst = ams.left_lists(al, bl);
                if (lst.empty()) continue;
                int max_ar = (int)al, max_br = (int)bl;
                max_right_ends((int)al, (int)bl, &max_ar, &max_br);
                forward_group(al, max_ar, bl, max_br, f);
                Mbar.resize(lenA + 1, lenB + 1, 0.0);
                Mbar2.resize(lenA + 1, lenB + 1, 0.0);
                bool any_sig = false;
                for (int x : lst) {
                    const ArcMatch &am = ams.ams[x];
                    const double seed = Dbar[x] * eam[x];
                    Mbar(am.a.right - 1, am.b.right - 1) += seed;
                    const double pam = Dmat(am.a.idx, am.b.idx) * Dbar[x] / Z;
                    if (pam > thr) {
                        Mbar2(am.a.right - 1, am.b.right - 1) += seed;
                        any_sig = true;
                    }
                }
                reverse_group(f, Mbar, &Dbar, &edge_full, &openings);
                if (any_sig) reverse_group(f, Mbar2, nullptr, &edge_ref, nullptr);
            }
        }
    }
};

}  // namespace orc

// ================================================================ C API
struct oracle_result {
    std::unique_ptr<orc::Inputs> in;
    long long score = 0;
    double pf = 0;
    std::vector<std::pair<int, int>> edges;
    std::string sA, sB;
    std::vector<long long> D;
    std::vector<long long> am_st;  // stacked arc-match score, or NEG_INF if not stackable
    std::vector<double> am_prob, bm_prob, bm_full, pfD, pfDprime;
    double openings = 0;
    long long counters[6] = {0, 0, 0, 0, 0, 0};
};

static void set_err(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg, errlen - 1);
        err[errlen - 1] = 0;
    }
}

extern "C" oracle_result *oracle_align(const char *ppA, const char *ppB, const oracle_params *p,
                                       const char *ribosum_path, int do_trace, char *err, int errlen) {
    try {
        auto *r = new oracle_result();
        r->in = orc::prepare(ppA, ppB, *p, ribosum_path);
        orc::Aligner al(*r->in);
        al.counting = true;
        r->score = al.align();
        r->counters[0] = (long long)r->in->bpsA->arcs.size();
        r->counters[1] = (long long)r->in->bpsB->arcs.size();
        r->counters[2] = (long long)r->in->am->ams.size();
        r->counters[3] = al.cnt_G;
        r->counters[4] = al.cnt_C;
        r->counters[5] = al.cnt_K;
        al.counting = false;
        r->D.resize(r->in->am->ams.size());
        for (size_t k = 0; k < r->D.size(); ++k) r->D[k] = al.D(r->in->am->ams[k]);
        r->am_st.assign(r->D.size(), orc::NEG_INF);
        for (size_t k = 0; k < r->D.size(); ++k)
            if (al.am_stackable[k]) r->am_st[k] = al.am_score_st[k];
        if (do_trace) {
            al.trace();
            r->edges = al.edges;
            r->sA = al.structA;
            r->sB = al.structB;
        }
        return r;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

// Penalized (mode 1, arg = position penalty) and normalized (mode 2, arg =
// L; sequence-local) alignment: score (normalized: the normalized score)
// and the traced alignment.
extern "C" oracle_result *oracle_align_variant(const char *ppA, const char *ppB, const oracle_params *p,
                                               const char *ribosum_path, int mode, long long arg, char *err,
                                               int errlen) {
    try {
        auto *r = new oracle_result();
        r->in = orc::prepare(ppA, ppB, *p, ribosum_path);
        orc::Aligner al(*r->in);
        if (mode == 1)
            r->score = al.penalized_align(arg);
        else if (mode == 2)
            r->score = al.normalized_align(arg);
        else
            throw std::runtime_error("mode: 1 penalized, 2 normalized");
        r->edges = al.edges;
        r->sA = al.structA;
        r->sB = al.structB;
        return r;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

// The band step's trace probabilities (PFTraceProbs, edge_probs.icc:212-268)
// over the full matrix, with the similarity sigma(i, j) (StralScore) of
// every cell and the Boltzmann gap weights, for the check of the Gotoh
// partition function against an enumeration of all alignments.
// probs, sim: (lenA+1) x (lenB+1) row major; w[0] = exp(gap opening / T),
// w[1] = exp(gap extension / T), w[2] = T, w[3] = lenA, w[4] = lenB.
// Returns the cell count.
extern "C" int oracle_trace_probs(const char *ppA, const char *ppB, const oracle_params *p,
                                  const char *ribosum_path, double *probs, double *sim, double *w, int cap) {
    try {
        using namespace orc;
        std::unique_ptr<Ribosum> rib;
        if (p->use_ribosum) rib.reset(new Ribosum(read_ribosum(ribosum_path)));
        RnaData A = read_pp(ppA, p->min_prob, p->max_bp_span), B = read_pp(ppB, p->min_prob, p->max_bp_span);
        const int n = A.length(), m = B.length();
        if (cap < (n + 1) * (m + 1)) return -1;
        double simm[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) simm[i][j] = rib ? rib->bm[i][j] : (i == j ? p->match : p->mismatch);
        Band band(n, m);
        if (p->max_diff >= 0) band.constrain_wo_ref((size_t)p->max_diff);
        const double go = p->indel_opening / 100.0, ge = p->indel / 100.0, sw = p->struct_weight / 100.0;
        const double T = p->temperature_alipf / 100.0;
        Mat<double> pr = p->band_extended
                             ? trace_probs<long double>(A, B, band, simm, go, ge, sw, T, p->sequ_local, p->free_endgaps)
                             : trace_probs<double>(A, B, band, simm, go, ge, sw, T, p->sequ_local, p->free_endgaps);
        StralScore sc(A, B, simm, sw, go, ge);
        for (int i = 0; i <= n; ++i)
            for (int j = 0; j <= m; ++j) {
                probs[i * (m + 1) + j] = pr(i, j);
                sim[i * (m + 1) + j] = (i > 0 && j > 0) ? sc.sigma(i, j) : 0.0;
            }
        w[0] = exp(go / T);
        w[1] = exp(ge / T);
        w[2] = T;
        w[3] = n;
        w[4] = m;
        return (n + 1) * (m + 1);
    } catch (std::exception &) {
        return -2;
    }
}

extern "C" void oracle_free(oracle_result *r) { delete r; }
extern "C" long long oracle_score(const oracle_result *r) { return r->score; }
extern "C" double oracle_partition_function(const oracle_result *r) { return r->pf; }
extern "C" int oracle_lengths(const oracle_result *r, int *a, int *b) {
    *a = r->in->A.length();
    *b = r->in->B.length();
    return 0;
}
extern "C" int oracle_edges(const oracle_result *r, int *a, int *b, int cap) {
    int n = (int)r->edges.size();
    for (int k = 0; k < n && k < cap; ++k) {
        a[k] = r->edges[k].first;
        b[k] = r->edges[k].second;
    }
    return n;
}
extern "C" int oracle_structures(const oracle_result *r, char *sa, char *sb, int cap) {
    if ((int)r->sA.size() >= cap || (int)r->sB.size() >= cap) return -1;
    std::strcpy(sa, r->sA.c_str());
    std::strcpy(sb, r->sB.c_str());
    return 0;
}
extern "C" int oracle_band(const oracle_result *r, int *mn, int *mx, int cap) {
    const auto &b = *r->in->band;
    int n = (int)b.min_col.size();
    for (int i = 0; i < n && i < cap; ++i) {
        mn[i] = (int)b.min_col[i];
        mx[i] = (int)b.max_col[i];
    }
    return n;
}
// stacked arc-match scores (--stacking) in arc-match order; not stackable:
// stackable[k] = 0
extern "C" int oracle_stacked_scores(const oracle_result *r, long long *score, int *stackable, int cap) {
    const int n = (int)r->am_st.size();
    if (cap < n) return -1;
    for (int k = 0; k < n; ++k) {
        stackable[k] = r->am_st[k] != orc::NEG_INF;
        score[k] = stackable[k] ? r->am_st[k] : 0;
    }
    return n;
}

extern "C" int oracle_arcmatches(const oracle_result *r, int *al, int *ar, int *bl, int *br, long long *score,
                                 long long *D, int cap) {
    const auto &ams = r->in->am->ams;
    int n = (int)ams.size();
    for (int k = 0; k < n && k < cap; ++k) {
        al[k] = ams[k].a.left;
        ar[k] = ams[k].a.right;
        bl[k] = ams[k].b.left;
        br[k] = ams[k].b.right;
        score[k] = r->in->sc->arcmatch(ams[k]);
        D[k] = k < (int)r->D.size() ? r->D[k] : 0;
    }
    return n;
}
extern "C" int oracle_tables(const oracle_result *r, long long *sigma, long long *gA, long long *gB) {
    const auto &sc = *r->in->sc;
    int la = r->in->A.length(), lb = r->in->B.length();
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) sigma[i * (lb + 1) + j] = sc.sigma_tab(i, j);
    for (int i = 0; i <= la; ++i) gA[i] = sc.gapA_tab[i];
    for (int j = 0; j <= lb; ++j) gB[j] = sc.gapB_tab[j];
    return 0;
}
extern "C" int oracle_counters(const oracle_result *r, long long *out6) {
    for (int k = 0; k < 6; ++k) out6[k] = r->counters[k];
    return 0;
}
extern "C" int oracle_am_probs(const oracle_result *r, double *prob, int cap) {
    int n = (int)r->am_prob.size();
    for (int k = 0; k < n && k < cap; ++k) prob[k] = r->am_prob[k];
    return n;
}
extern "C" int oracle_bm_probs(const oracle_result *r, double *prob, int cap) {
    int n = (int)r->bm_prob.size();
    for (int k = 0; k < n && k < cap; ++k) prob[k] = r->bm_prob[k];
    return n;
}
extern "C" int oracle_pf_D(const oracle_result *r, double *D, double *Dp, int cap) {
    int n = (int)r->pfD.size();
    for (int k = 0; k < n && k < cap; ++k) {
        D[k] = r->pfD[k];
        Dp[k] = k < (int)r->pfDprime.size() ? r->pfDprime[k] : 0.0;
    }
    return n;
}

extern "C" int oracle_align_many(const char **ppA, const char **ppB, int npairs, const oracle_params *p,
                                 const char *ribosum_path, int nthreads, long long *scores) {
    std::atomic<int> next(0), failed(0);
    auto work = [&]() {
        for (;;) {
            int k = next.fetch_add(1);
            if (k >= npairs) break;
            try {
                auto in = orc::prepare(ppA[k], ppB[k], *p, ribosum_path);
                orc::Aligner al(*in);
                scores[k] = al.align();
                al.trace();
            } catch (...) {
                failed++;
            }
        }
    };
    std::vector<std::thread> th;
    if (nthreads < 1) nthreads = 1;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
    for (auto &t : th) t.join();
    return failed ? -1 : npairs;
}

// k-best suboptimal alignments (Aligner::suboptimal): scores[t] and, per
// alignment, edge_count[t] edges at edge_a / edge_b + t * cap_edges.
// Returns the number of alignments or -1 (message in err).
extern "C" int oracle_suboptimal(const char *ppA, const char *ppB, const oracle_params *p, const char *ribosum_path,
                                 int k, long long threshold, long long *scores, int *restr, int *edge_a,
                                 int *edge_b, int *edge_count, int cap_alns, int cap_edges, char *err,
                                 int errlen) {
    try {
        auto in = orc::prepare(ppA, ppB, *p, ribosum_path);
        orc::Aligner al(*in);
        const auto res = al.suboptimal(k, threshold);
        const int n = (int)std::min<size_t>(res.size(), (size_t)cap_alns);
        for (int t = 0; t < n; ++t) {
            scores[t] = res[t].score;
            if (restr) {
                restr[4 * t] = res[t].sA;
                restr[4 * t + 1] = res[t].eA;
                restr[4 * t + 2] = res[t].sB;
                restr[4 * t + 3] = res[t].eB;
            }
            const int ne = (int)std::min<size_t>(res[t].edges.size(), (size_t)cap_edges);
            edge_count[t] = ne;
            for (int e = 0; e < ne; ++e) {
                edge_a[(size_t)t * cap_edges + e] = res[t].edges[e].first;
                edge_b[(size_t)t * cap_edges + e] = res[t].edges[e].second;
            }
        }
        return n;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return -1;
    }
}

// fragment match probabilities of the LocARNA-P problem of r (an
// oracle_pf result): prob[q] for the fragments ijkl[4q .. 4q+3] = i, j, k, l
extern "C" int oracle_pf_fragments(oracle_result *r, int nq, const int *ijkl, double *prob, char *err, int errlen) {
    try {
        if (!r || nq < 0 || (nq > 0 && (!ijkl || !prob))) throw std::runtime_error("bad arguments");
        orc::AlignerP ap(*r->in);
        ap.inside();
        std::vector<double> Dbar;
        orc::Mat<double> edge_full, edge_ref;
        double openings = 0;
        ap.outside(Dbar, edge_full, edge_ref, openings);
        for (int q = 0; q < nq; ++q) {
            const int *x = ijkl + 4 * q;
            if (x[0] < 1 || x[1] < x[0] || x[2] < 1 || x[3] < x[2]) throw std::runtime_error("fragment positions out of range");
            prob[q] = ap.fragment(Dbar, (size_t)x[0], (size_t)x[1], (size_t)x[2], (size_t)x[3]);
        }
        return nq;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return -1;
    }
}

extern "C" oracle_result *oracle_pf(const char *ppA, const char *ppB, const oracle_params *p,
                                    const char *ribosum_path, int include_am_in_bm, char *err, int errlen) {
    try {
        if (p->free_endgaps) throw std::runtime_error("free end gaps: integer path only");
        auto *r = new oracle_result();
        r->in = orc::prepare(ppA, ppB, *p, ribosum_path);
        orc::AlignerP ap(*r->in);
        r->pf = ap.inside();
        r->counters[0] = (long long)r->in->bpsA->arcs.size();
        r->counters[1] = (long long)r->in->bpsB->arcs.size();
        r->counters[2] = (long long)r->in->am->ams.size();
        std::vector<double> Dbar;
        orc::Mat<double> edge_full, edge_ref;
        double openings = 0;
        ap.outside(Dbar, edge_full, edge_ref, openings);
        const auto &ams = r->in->am->ams;
        const int la = r->in->A.length(), lb = r->in->B.length();
        r->am_prob.resize(ams.size());
        r->pfD.resize(ams.size());
        r->pfDprime.resize(ams.size());
        for (size_t k = 0; k < ams.size(); ++k) {
            const double D = ap.Dmat(ams[k].a.idx, ams[k].b.idx);
            r->pfD[k] = D;
            r->pfDprime[k] = Dbar[k] * ap.eam[k] / ap.ps;  // the reference's D'
            r->am_prob[k] = D * Dbar[k] / ap.Z;
        }
        r->bm_prob.assign((size_t)(la + 1) * (lb + 1), 0.0);
        r->bm_full.assign((size_t)(la + 1) * (lb + 1), 0.0);
        const auto &tc = *r->in->band;
        for (int i = 1; i <= la; ++i)
            for (int j = 1; j <= lb; ++j) {
                r->bm_full[(size_t)i * (lb + 1) + j] = edge_full(i, j) / ap.Z;
                if (tc.is_valid(i, j) && tc.is_valid_match(i, j))
                    r->bm_prob[(size_t)i * (lb + 1) + j] = edge_ref(i, j) / ap.Z;
            }
        if (include_am_in_bm) {
            for (size_t k = 0; k < ams.size(); ++k) {
                r->bm_prob[(size_t)ams[k].a.left * (lb + 1) + ams[k].b.left] += r->am_prob[k];
                r->bm_prob[(size_t)ams[k].a.right * (lb + 1) + ams[k].b.right] += r->am_prob[k];
            }
        }
        r->openings = openings / ap.Z;
        return r;
    } catch (std::exception &e) {
        set_err(err, errlen, e.what());
        return nullptr;
    }
}

extern "C" int oracle_pf_extra(const oracle_result *r, double *bm_full, int cap, double *openings) {
    int n = (int)r->bm_full.size();
    for (int k = 0; k < n && k < cap; ++k) bm_full[k] = r->bm_full[k];
    if (openings) *openings = r->openings;
    return n;
}

extern "C" long long oracle_consensus_pp(const char *ppA, const char *ppB, const oracle_params *p, const int *ea,
                                         const int *eb, int nedges, char *buf, long long cap, char *err, int errlen) {
    try {
        const orc::RnaData A = orc::read_pp(ppA, p->min_prob, p->max_bp_span, p->stacking != 0);
        const orc::RnaData B = orc::read_pp(ppB, p->min_prob, p->max_bp_span, p->stacking != 0);
        std::vector<std::pair<int, int>> edges;
        for (int k = 0; k < nedges; ++k) {
            if ((ea[k] < 0 && eb[k] < 0) || ea[k] > A.length() || eb[k] > B.length() || ea[k] == 0 || eb[k] == 0)
                throw orc::failure("invalid alignment edge");
            edges.push_back({ea[k], eb[k]});
        }
        // locarna.cc:662-663
        const double peA = p->exp_prob > 0 ? p->exp_prob : 1.0 / (2.0 * A.length());
        const double peB = p->exp_prob > 0 ? p->exp_prob : 1.0 / (2.0 * B.length());
        const std::string text = orc::consensus_pp(A, B, edges, peA, peB, p->stacking != 0);
        if ((long long)text.size() + 1 <= cap) std::memcpy(buf, text.c_str(), text.size() + 1);
        return (long long)text.size();
    } catch (const std::exception &e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return -1;
    }
}

extern "C" int oracle_format_prob(double prob, char *buf, int cap) {
    const std::string s = orc::format_prob(prob);
    if ((int)s.size() + 1 > cap) return -1;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return (int)s.size();
}
