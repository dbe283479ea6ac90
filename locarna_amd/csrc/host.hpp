// Host side of the MI355X LocARNA aligner: input data, score tables, trace
// band and arc-match lists, in the shape the gfx950 kernels consume.
//
// The classes keep the names and meaning of the reference interface
// (LocARNA 2.0.1, src/LocARNA/*.hh) so that callers of RnaData, BasePairs,
// TraceController, ArcMatches, Scoring and Aligner find the same objects;
// the data layout is our own (flat arrays, CSR lists, int32 scores).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace locarna_amd {

typedef long score_t;  // reference: scoring_fwd.hh:13-17

// reference: aux.hh:229-278 (LocARNA::failure)
struct failure : std::runtime_error {
    explicit failure(const std::string &s) : std::runtime_error(s) {}
};
// A well-formed problem beyond what the kernels represent (LCG_E_CAPACITY).
struct capacity_failure : failure {
    explicit capacity_failure(const std::string &s) : failure(s) {}
};

// Sequence, or multiple alignment (a profile: mlocarna's intermediate
// alignments), plus filtered base-pair probabilities read from #PP 2.0 text.
// reference: RnaData, rna_data.hh:37; reader rna_data.cc:984-1124, rows
// multiple_alignment.cc:270-380
class RnaData {
public:
    // stacking: read the joint probabilities of the #STACK column
    RnaData(const std::string &pp_text, double p_bpcut, long max_bp_span, bool stacking = false);
    int length() const { return (int)seq_.size(); }
    const std::string &sequence() const { return seq_; }  // first row, 0-based chars
    char nt(int i) const { return seq_[i - 1]; }           // first row, 1-based
    const std::string &name() const { return name_; }
    // rows of the alignment, in first-seen name order (one for a sequence)
    int num_rows() const { return (int)rows_.size(); }
    const std::string &row(int k) const { return rows_[k]; }
    const std::string &row_name(int k) const { return names_[k]; }
    double cutoff() const { return p_bpcut_; }  // arc_cutoff_prob (rna_data.cc:686-688)
    double arc_prob(int i, int j) const;
    // joint probability of (i, j) and (i+1, j-1); 0 without (rna_data.cc:700-710)
    double joint_arc_prob(int i, int j) const;
    double prob_paired_upstream(int i) const;    // rna_data.cc:713-722
    double prob_paired_downstream(int i) const;  // rna_data.cc:724-732
    // base pairs kept after filtering, ordered i ascending, j ascending
    const std::vector<std::pair<std::pair<int, int>, double>> &pairs() const { return pairs_; }

private:
    std::string name_, seq_;
    std::vector<std::string> names_, rows_;
    std::map<std::pair<int, int>, double> probs_, joint_;
    std::vector<double> up_, down_;  // prob_paired_upstream / downstream per position
    std::vector<std::pair<std::pair<int, int>, double>> pairs_;
    double p_bpcut_;
};

struct Arc {
    int idx, left, right;
};

// reference: BasePairs, basepairs.hh; basepairs.cc:156-220
class BasePairs {
public:
    BasePairs(const RnaData *rna, double min_prob);
    // arcs given as a set, registered in (left, right) ascending order
    // (BasePairs(len, bpair_set), basepairs.cc:79-94)
    BasePairs(int len, const std::vector<std::pair<int, int>> &sorted_arcs);
    // index of arc (i, j), -1 if absent
    int find(int i, int j) const;
    int num_bps() const { return (int)arcs_.size(); }
    const Arc &arc(int idx) const { return arcs_[idx]; }
    int seqlen() const { return len_; }
    // arcs with right end j sorted by left end descending (right_adjlist)
    const std::vector<int> &right_adj(int j) const { return right_[j]; }
    // arcs with left end i sorted by right end ascending (left_adjlist)
    const std::vector<int> &left_adj(int i) const { return left_[i]; }

private:
    int len_;
    std::vector<Arc> arcs_;
    std::vector<std::vector<int>> left_, right_;
};

// Per-row valid column range of the DP matrices.
// reference: TraceController, trace_controller.hh:328-336
class TraceController {
public:
    TraceController(int lenA, int lenB);
    int rows() const { return lenA_; }
    int min_col(int i) const { return min_col_[i]; }
    int max_col(int i) const { return max_col_[i]; }
    bool is_valid(int i, int j) const { return min_col_[i] <= j && j <= max_col_[i]; }
    bool is_valid_match(int i, int j) const { return is_valid(i, j) && is_valid(i - 1, j - 1); }
    // trace_controller.cc:565-597
    void restrict_by_trace_probabilities(const std::vector<double> &probs, double min_prob);
    // band of width max-diff around the diagonal without a reference
    // alignment (constrain_wo_ref, trace_controller.cc:361-390)
    void constrain_wo_ref(int delta);
    // the same from the first and last column of each row whose trace
    // probability reaches the cut (-1: no column of the row does)
    void restrict_by_rows(const std::vector<int> &first, const std::vector<int> &last);
    const std::vector<int> &min_cols() const { return min_col_; }
    const std::vector<int> &max_cols() const { return max_col_; }

private:
    int lenA_, lenB_;
    std::vector<int> min_col_, max_col_;
};

// Numeric parameters; names follow the locarna / locarna_p command lines
// (src/locarna.cc:83-272, src/locarna_p.cc:95-200) and ScoringParams
// (scoring.hh:51-225).
struct ScoringParams {
    int indel = -150;
    int indel_opening = -750;
    int struct_weight = 200;
    int tau = 50;
    int match = 50;
    int mismatch = 0;
    int temperature_alipf = 300;
    bool use_ribosum = true;
    double exp_prob = -1.0;  // <= 0: 1/(2 len)  (aux.hh:216-220)
    int stacking = 0;        // bit 0 --stacking, bit 1 --new-stacking
};

// RIBOSUM nucleotide code: 0..3 for A, C, G, U; -1 for any other symbol
int nt_code(char c);

// Base similarity of the trace band's Gotoh partition function: RIBOSUM
// base-match scores, or match / mismatch (main_helper.icc:284-308).
void band_similarity(const ScoringParams &sp, double sim[16]);

// Trace-probability band: Gotoh partition function over the StralScore
// similarity, forward and backward (free end gaps: LCG_FREE_* bits), then
// the probability cut and the monotone fix-ups.  reference:
// edge_probs.icc:7-162, :212-268;
// stral_score.cc:29-60; main_helper.icc:371-426
void restrict_trace_by_probabilities(TraceController &tc, const RnaData &A, const RnaData &B,
                                     const ScoringParams &sp, double min_trace_prob,
                                     bool extended_precision, bool local, int free_endgaps = 0);

struct ArcMatch {
    int idx;
    int arcA, arcB;  // arc indices
};

// One line of an arc-match scores file: arcs (i, j) of A and (k, l) of B
// and the score of matching them.
struct ExplicitArcMatch {
    int i, j, k, l;
    score_t score;
};
// Parses an arc-match scores text ("i j k l score" per line; with
// probability_scale >= 0 the fifth field is a probability p and the score
// (long)(p * probability_scale)).  reference:
// ArcMatches::read_arcmatch_scores, arc_matches.cc:190-246.  Throws failure
// on a malformed line, on i > j, k > l, a zero or too large position (as
// the reference), and on arcs with j - i < 3 or l - k < 3, which the
// reference accepts but its own BasePairs never registers
// (basepairs.cc:162-177) and whose D would read outside their group.
std::vector<ExplicitArcMatch> parse_arcmatch_scores(const std::string &text, int lenA, int lenB,
                                                    int probability_scale);

// Valid arc matches in reference order (arc A index outer, arc B inner).
// reference: ArcMatches, arc_matches.cc:19-48, 130-188
class ArcMatches {
public:
    ArcMatches(const BasePairs &A, const BasePairs &B, const TraceController &tc, int max_diff_am = -1,
               int max_diff_at_am = -1);
    // no arc matches enumerated on the host: the device compaction builds
    // the lists (the object still carries the two base-pair sets)
    ArcMatches(const BasePairs &A, const BasePairs &B) : A_(A), B_(B) {}
    // arc matches of an arc-match scores file, in line order, the invalid
    // ones left out, with their explicit scores (arc_matches.cc:248-273)
    ArcMatches(const BasePairs &A, const BasePairs &B, const TraceController &tc,
               const std::vector<ExplicitArcMatch> &lines, int max_diff_am = -1, int max_diff_at_am = -1);
    bool explicit_scores() const { return explicit_; }
    score_t explicit_score(int k) const { return scores_[k]; }
    int num_arc_matches() const { return (int)ams_.size(); }
    const ArcMatch &arcmatch(int k) const { return ams_[k]; }
    const BasePairs &bpsA() const { return A_; }
    const BasePairs &bpsB() const { return B_; }

private:
    const BasePairs &A_, &B_;
    std::vector<ArcMatch> ams_;
    bool explicit_ = false;
    std::vector<score_t> scores_;
};

// Integer score tables.  reference: Scoring, scoring.cc:28-554
class Scoring {
public:
    Scoring(const RnaData &A, const RnaData &B, const ArcMatches &am, const ScoringParams &p);
    static score_t round2score(double d) { return (score_t)((d < 0) ? (d - 0.5) : (d + 0.5)); }
    score_t basematch(int i, int j) const;  // sigma(i, j)
    score_t gapA(int i) const { return gapA_[i]; }
    score_t gapB(int j) const { return gapB_[j]; }
    score_t indel_opening() const { return p_.indel_opening; }
    score_t arcmatch(int am) const { return am_score_[am]; }
    // --stacking: whether both arcs of the arc match have a joint
    // probability with their inner arcs (scoring.cc:556-569), and its
    // stacked score (scoring.cc:440-554 with the stacked weights of
    // scoring.cc:200-236)
    bool stackable(int am) const { return !am_stackable_.empty() && am_stackable_[am]; }
    score_t arcmatch_stacked(int am) const { return am_score_st_[am]; }
    score_t weightA(int arc) const { return wA_[arc]; }
    score_t weightB(int arc) const { return wB_[arc]; }
    const ScoringParams &params() const { return p_; }
    // a pair of profiles (more than one row on a side): sigma as a table
    // over all column pairs instead of symbol codes
    bool profile() const { return profile_; }
    // single-sequence symbol code (0..3 = ACGU, 4.. other symbols; 0 for
    // profiles)
    int codeA(int i) const { return codeA_[i]; }
    int codeB(int j) const { return codeB_[j]; }
    static const int NCODES = 8;
    score_t sigma_code(int a, int b) const { return sigma_codes_[a * NCODES + b]; }
    // largest |sigma| over the code table (sequences) or the column table
    // (profiles)
    score_t max_abs_sigma() const;
    // profiles: sigma of columns (i, j), 1-based
    const std::vector<score_t> &sigma_table() const { return sigma_tab_; }
    // ribosum arc-match term riboX (scoring.cc:368-438) before the tau factor
    score_t arc_sequence_term(int am) const { return am_seq_[am]; }
    // riboX term of one arc match from the nucleotide codes (0..3 = ACGU) of
    // its four ends, RIBOSUM85_60 (scoring.cc:368-438, single sequences)
    static score_t ribosum_arc_term(int al, int ar, int bl, int br);
    // the same averaged over the row pairs of two profiles whose four ends
    // are all in ACGU (scoring.cc:368-438)
    static score_t ribosum_arc_term_rows(const RnaData &A, const Arc &a, const RnaData &B, const Arc &b);

private:
    ScoringParams p_;
    const ArcMatches &am_;
    bool profile_ = false;
    int lenB_ = 0;
    std::vector<score_t> sigma_tab_;  // profiles: (n+1) x (m+1)
    std::vector<int> codeA_, codeB_;
    std::vector<char> symbols_;
    std::vector<score_t> sigma_codes_;
    std::vector<score_t> gapA_, gapB_, wA_, wB_, am_score_, am_seq_, am_score_st_;
    std::vector<char> am_stackable_;
};

// ------------------------------------------------------------ writers
// Rows of an aligned pair: the rows of A, then those of B, with '-' in the
// gap columns of `cols` ((a, b) per column, -1 for a gap); on a name clash
// between the two sides the names get "A." / "B." prefixes.
// reference: MultipleAlignment::init, multiple_alignment.cc:186-247
std::vector<std::pair<std::string, std::string>> aligned_rows(const RnaData &A, const RnaData &B,
                                                              const std::vector<std::pair<int, int>> &cols);
// Name / sequence lines in blocks of `width` columns, names padded to
// max(18, longest name) and one blank, a blank line between blocks.
// reference: MultipleAlignment::write, multiple_alignment.cc:994-1085
std::string clustal_blocks(const std::vector<std::pair<std::string, const std::string *>> &rows, size_t width);
// A probability as pp files write it: 4 significant digits, or scientific
// with 3 decimals when that is shorter, "e-0" as "e-".
// reference: format_prob, rna_data.cc:1274-1303
std::string format_prob(double p);
// pp 2.0 text of an aligned pair (locarna --pp): the aligned rows, and the
// consensus base-pair probabilities of every pair of columns as weighted
// geometric means of the two sides' probabilities, each raised to at
// least min(expected probability, 0.1 x cut), kept above the weighted
// geometric mean of the two cuts, which is written as #BPCUT.  With
// stacking (both sides read with joint probabilities) a pair is kept also
// when the consensus of the joint probabilities exceeds the cut; the text
// carries no joint column (the reference's consensus object has no
// stacking flag).  Lines of base pairs in (i, j) order.
// reference: RnaData consensus constructor, rna_data.cc:75-122,
// init_as_consensus_dot_plot / consensus_probability :1475-1570, write_pp
// :1242-1339
std::string consensus_pp_text(const RnaData &A, const RnaData &B, const std::vector<std::pair<int, int>> &cols,
                              double p_expA, double p_expB, bool stacking);

}  // namespace locarna_amd
