// C++ facade of the MI355X LocARNA alignment library, shaped after the
// reference's in-process API so that its callers (src/locarna.cc:727-785,
// src/locarna_p.cc:466-561) change their includes and object construction,
// not their control flow:
//
//   LocARNA                                     here
//   Aligner(const AlignerParams &)              Aligner(const AlignerParams &)       aligner.hh:94
//     align(), trace(), get_alignment()           same                               aligner.hh:101-120
//     normalized_align(L), penalized_align(P)     same                               aligner.hh:166-171
//     suboptimal(k, threshold, ...)               suboptimal(k, threshold) -> alignments aligner.hh:146
//   AlignerP<T>(const AlignerPParams<T> &)      AlignerP<double>                      aligner_p.hh:501
//     align_inside(), align_outside()             same                               aligner_p.hh:528-536
//     compute_arcmatch_probabilities()            same                               aligner_p.hh:546
//     compute_basematch_probabilities(bool)       same                               aligner_p.hh:541
//     write_arcmatch_probabilities(ostream &)     same                               aligner_p.hh:555
//     write_basematch_probabilities(ostream &)    same                               aligner_p.hh:564
//     compute_fragment_match_prob(i, j, k, l)     same                               aligner_p.hh:605
//   LocARNA::failure (aux.hh:229-255)           locarna_amd::failure (status code kept)
//
// Parameters use the reference's named-argument style
// (aligner_params.hh:244-349): AlignerParams(AlignerParams::seqA(&a),
// AlignerParams::seqB(&b), AlignerParams::scoring(&scoring_params), ...).
// Where the reference passes objects it has built itself (Sequence,
// Scoring, ArcMatches, TraceController), this facade passes what they are
// built from -- the #PP 2.0 texts (RnaData), the ScoringParams fields and
// the band options -- and the library builds them (lcg_pair_prepare).
// Features the library does not implement (anchor constraints;
// no_lonely_pairs, stacking and struct_local for AlignerP) fail with
// failure(LCG_E_ARG) when asked for.
//
// Header only, over the C ABI of locarna_amd.h; link -llocarna_amd.
#ifndef LOCARNA_AMD_HPP
#define LOCARNA_AMD_HPP

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <utility>
#include <vector>

#include "locarna_amd.h"

namespace locarna_amd {

// Error of any library call: the reference throws LocARNA::failure
// (aux.hh:229-255); the status code of the C ABI is kept in code().
class failure : public std::runtime_error {
public:
    failure(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

namespace detail {
inline void check(int rc, const char *what, const char *err = nullptr) {
    if (rc >= 0) return;
    std::string msg = std::string(what) + " failed (status " + std::to_string(rc) + ")";
    if (err && err[0]) msg += ": " + std::string(err);
    throw failure(rc, msg);
}
}  // namespace detail

// Minus infinity of a score (the reference's InftyInt -inf, infty_int.cc:7-23):
// every value below -2^28 means minus infinity.
inline bool is_neg_infty(int64_t score) { return score < -(int64_t(1) << 28); }

// Sequence with base-pair probabilities, from #PP 2.0 text (RnaData,
// rna_data.hh:37; reader rna_data.cc:984-1124).
class RnaData {
public:
    explicit RnaData(std::string pp_text) : pp_(std::move(pp_text)) {}
    static RnaData from_file(const std::string &path) {
        std::ifstream in(path);
        if (!in) throw failure(LCG_E_INPUT, "Cannot read file " + path);
        std::ostringstream s;
        s << in.rdbuf();
        return RnaData(s.str());
    }
    const std::string &pp_text() const { return pp_; }

private:
    std::string pp_;
};

// Score parameters, named as ScoringParams (scoring.hh:82-260), with the
// defaults of the locarna / locarna_p options that fill it
// (src/locarna.cc:99-217, src/locarna_p.cc:101-119).
struct ScoringParams {
    int match = 50;
    int mismatch = 0;
    int indel = -150;
    int indel_opening = -750;
    bool use_ribosum = true;  // built-in RIBOSUM85_60
    int struct_weight = 200;
    int tau_factor = 50;
    double exp_prob = -1.0;  // <= 0: 1 / (2 length)
    int temperature_alipf = 300;
    // stacking terms (scoring.hh:116-120): stacking derives stacked arc
    // weights from the joint probabilities of #STACK files, new_stacking
    // adds them to the plain weights; the DP follows these flags
    // (aligner.cc:600, :629, :987)
    bool stacking = false;
    bool new_stacking = false;
    int exclusion = 0;  // score of one excluded block with struct_local (scoring.hh:107)
};

// Base-pair and band options (the RnaData / TraceController arguments of
// src/locarna.cc:465-590; defaults :183-200, :253).
struct BandParams {
    double min_prob = 0.001;                // --min-prob
    int max_bp_span = -1;                   // --maxBPspan (-1: unlimited)
    double min_trace_probability = 1e-4;    // --min-trace-probability (<= 0: no band)
    int max_diff = -1;                      // --max-diff (-1: none)
    bool extended_precision_band = true;    // locarna: long double; locarna_p: double
};

// Named-argument tags (the reference's DEFINE_NAMED_ARG_FEATURE).
#define LCG_NAMED_ARG(name, type)                        \
    struct name {                                        \
        explicit name(type v) : value(std::move(v)) {}  \
        type value;                                      \
    };

class AlignerParams {
public:
    LCG_NAMED_ARG(seqA, const RnaData *)
    LCG_NAMED_ARG(seqB, const RnaData *)
    LCG_NAMED_ARG(scoring, const ScoringParams *)
    LCG_NAMED_ARG(trace_controller, const BandParams *)
    LCG_NAMED_ARG(no_lonely_pairs, bool)
    LCG_NAMED_ARG(struct_local, bool)
    LCG_NAMED_ARG(sequ_local, bool)
    LCG_NAMED_ARG(free_endgaps, std::string)  // "----" as FreeEndgaps (free_endgaps.hh)
    LCG_NAMED_ARG(DO_TRACE, bool)
    LCG_NAMED_ARG(max_diff_am, int)
    LCG_NAMED_ARG(max_diff_at_am, int)
    LCG_NAMED_ARG(stacking, bool)
    LCG_NAMED_ARG(device, int)  // HIP device ordinal (no counterpart in the reference)

    template <class... Args>
    explicit AlignerParams(Args... args) {
        (set(args), ...);
    }

    const RnaData *seqA_ = nullptr, *seqB_ = nullptr;
    ScoringParams scoring_;
    BandParams band_;
    bool sequ_local_ = false, do_trace_ = true, no_lonely_pairs_ = false, stacking_ = false, struct_local_ = false;
    std::string free_endgaps_ = "----";
    int max_diff_am_ = -1, max_diff_at_am_ = -1, device_ = 0;

    // the lcg_params of these parameters (tau, band precision and
    // temperature follow locarna unless set otherwise)
    lcg_params c_params() const {
        lcg_params p;
        std::memset(&p, 0, sizeof p);
        p.indel = scoring_.indel;
        p.indel_opening = scoring_.indel_opening;
        p.struct_weight = scoring_.struct_weight;
        p.tau = scoring_.tau_factor;
        p.match = scoring_.match;
        p.mismatch = scoring_.mismatch;
        p.temperature_alipf = scoring_.temperature_alipf;
        p.sequ_local = sequ_local_ ? 1 : 0;
        p.max_bp_span = band_.max_bp_span;
        p.band_extended = band_.extended_precision_band ? 1 : 0;
        p.use_ribosum = scoring_.use_ribosum ? 1 : 0;
        if (free_endgaps_.size() != 4) throw failure(LCG_E_ARG, "free end gaps: four characters, e.g. \"++--\"");
        for (int k = 0; k < 4; ++k)
            if (free_endgaps_[k] == '+') p.free_endgaps |= 1 << k;
        p.max_diff = band_.max_diff;
        p.max_diff_am = max_diff_am_;
        p.max_diff_at_am = max_diff_at_am_;
        p.min_prob = band_.min_prob;
        p.min_trace_prob = band_.min_trace_probability;
        p.exp_prob = scoring_.exp_prob;
        p.pf_scale = 1.0;
        p.min_am_prob = 0.0;
        p.min_bm_prob = 0.0;
        p.no_lonely_pairs = no_lonely_pairs_ ? 1 : 0;
        p.stacking = (scoring_.stacking ? 1 : 0) | (scoring_.new_stacking ? 2 : 0);
        p.struct_local = struct_local_ ? 1 : 0;
        p.exclusion = scoring_.exclusion;
        return p;
    }

protected:
    AlignerParams() = default;
    void set(const seqA &a) { seqA_ = a.value; }
    void set(const seqB &a) { seqB_ = a.value; }
    void set(const scoring &a) { scoring_ = *a.value; }
    void set(const trace_controller &a) { band_ = *a.value; }
    void set(const no_lonely_pairs &a) { no_lonely_pairs_ = a.value; }
    void set(const struct_local &a) { struct_local_ = a.value; }
    // kept as the reference keeps it (aligner_params.hh:63); the stacked
    // recursion follows ScoringParams::stacking / new_stacking
    void set(const stacking &a) { stacking_ = a.value; }
    void set(const sequ_local &a) { sequ_local_ = a.value; }
    void set(const free_endgaps &a) { free_endgaps_ = a.value; }
    void set(const DO_TRACE &a) { do_trace_ = a.value; }
    void set(const max_diff_am &a) { max_diff_am_ = a.value; }
    void set(const max_diff_at_am &a) { max_diff_at_am_ = a.value; }
    void set(const device &a) { device_ = a.value; }
    static void unsupported(bool on, const char *what) {
        if (on) throw failure(LCG_E_ARG, std::string(what) + " is not supported by this library");
    }
};

// LocARNA-P parameters (AlignerPParams<T>, aligner_params.hh:314-349);
// defaults of locarna_p (src/locarna_p.cc:95-200): double band,
// min-trace-probability 1e-5.
template <typename T = double>
class AlignerPParams : public AlignerParams {
    static_assert(std::is_same<T, double>::value, "the library computes partition functions in double");

public:
    LCG_NAMED_ARG(min_am_prob, double)
    LCG_NAMED_ARG(min_bm_prob, double)
    LCG_NAMED_ARG(pf_scale, double)

    template <class... Args>
    explicit AlignerPParams(Args... args) : AlignerParams() {
        band_.extended_precision_band = false;
        band_.min_trace_probability = 1e-5;
        (setp(args), ...);
    }

    double min_am_prob_ = 0.001, min_bm_prob_ = 0.001, pf_scale_ = 1.0;  // src/locarna_p.cc:124, 143-146

    lcg_params c_params() const {
        lcg_params p = AlignerParams::c_params();
        p.pf_scale = pf_scale_;
        p.min_am_prob = min_am_prob_;
        p.min_bm_prob = min_bm_prob_;
        return p;
    }

private:
    template <class A>
    void setp(const A &a) {
        set(a);
    }
    void setp(const min_am_prob &a) { min_am_prob_ = a.value; }
    void setp(const min_bm_prob &a) { min_bm_prob_ = a.value; }
    void setp(const pf_scale &a) { pf_scale_ = a.value; }
};
#undef LCG_NAMED_ARG

namespace detail {
struct PairDeleter {
    void operator()(lcg_pair *p) const { lcg_pair_free(p); }
};
struct BatchDeleter {
    void operator()(lcg_batch *b) const { lcg_batch_destroy(b); }
};
using PairPtr = std::unique_ptr<lcg_pair, PairDeleter>;
using BatchPtr = std::unique_ptr<lcg_batch, BatchDeleter>;

// one pair prepared on the host and uploaded as a one-pair batch
inline std::pair<PairPtr, BatchPtr> make_pair_batch(const AlignerParams &ap, const lcg_params &p) {
    if (!ap.seqA_ || !ap.seqB_) throw failure(LCG_E_ARG, "AlignerParams: seqA and seqB are required");
    char err[512] = {0};
    lcg_pair *pair = nullptr;
    check(lcg_pair_prepare(ap.seqA_->pp_text().c_str(), ap.seqB_->pp_text().c_str(), &p, &pair, err, sizeof err),
          "lcg_pair_prepare", err);
    PairPtr pp(pair);
    lcg_batch *batch = nullptr;
    check(lcg_batch_create(&pair, 1, ap.device_, &batch, err, sizeof err), "lcg_batch_create", err);
    return {std::move(pp), BatchPtr(batch)};
}

template <class F>
inline std::string text_of(F f) {
    const int64_t n = f(nullptr, 0);
    if (n < 0) check((int)n, "text writer");
    std::string s((size_t)n + 1, '\0');
    f(&s[0], n + 1);
    s.resize((size_t)n);
    return s;
}
}  // namespace detail

// Alignment of two sequences (Alignment, alignment.hh:83-280): edges in
// alignment order, positions 1-based, -1 for a gap.
class Alignment {
public:
    using edge_t = std::pair<int, int>;
    const std::vector<edge_t> &alignment_edges() const { return edges_; }
    const std::string &dot_bracket_structureA() const { return structA_; }
    const std::string &dot_bracket_structureB() const { return structB_; }
    int64_t score() const { return score_; }
    // what locarna --clustal writes (main_helper.icc:557-580)
    const std::string &clustal() const { return clustal_; }
    void write_clustal(std::ostream &out) const { out << clustal_; }

private:
    friend class Aligner;
    std::vector<edge_t> edges_;
    std::string structA_, structB_, clustal_;
    int64_t score_ = 0;
};

// Integer sequence-structure alignment of one pair (Aligner, aligner.hh:94-171).
class Aligner {
public:
    explicit Aligner(const AlignerParams &ap) : ap_(ap) {
        lcg_params p = ap.c_params();
        auto pb = detail::make_pair_batch(ap, p);
        pair_ = std::move(pb.first);
        batch_ = std::move(pb.second);
        detail::check(lcg_pair_info(pair_.get(), &lenA_, &lenB_, nullptr, nullptr, nullptr), "lcg_pair_info");
    }

    // Runs the stream-ordered launches on `stream` (a hipStream_t; nullptr:
    // the library's own); no counterpart in the reference.
    void set_stream(void *stream) { detail::check(lcg_batch_set_stream(batch_.get(), stream), "lcg_batch_set_stream"); }

    // Score of the best alignment (Aligner::align, aligner.cc:924-962);
    // minus infinity as -2^29 (is_neg_infty).
    int64_t align() {
        detail::check(lcg_batch_run(batch_.get(), 0, ap_.do_trace_ ? 1 : 0, nullptr), "lcg_batch_run");
        traced_ = ap_.do_trace_;
        return score();
    }
    // Traceback (Aligner::trace, aligner.cc:1345-1358); runs align() with
    // the traceback if that has not happened yet.
    void trace() {
        if (!traced_) {
            detail::check(lcg_batch_run(batch_.get(), 0, 1, nullptr), "lcg_batch_run");
            traced_ = true;
        }
        alignment_.reset();
    }
    // Penalized alignment (aligner.cc:1603-1622) with traceback.
    int64_t penalized_align(int64_t position_penalty) {
        detail::check(lcg_batch_run_penalized(batch_.get(), position_penalty, nullptr), "lcg_batch_run_penalized");
        traced_ = true;
        alignment_.reset();
        return score();
    }
    // Normalized alignment (aligner.cc:1526-1593; sequence-local only).
    int64_t normalized_align(int64_t L) {
        int32_t it = 0;
        detail::check(lcg_batch_run_normalized(batch_.get(), L, 100, &it, nullptr), "lcg_batch_run_normalized");
        traced_ = true;
        alignment_.reset();
        return score();
    }
    // k-best suboptimal alignments by interval splitting (Aligner::suboptimal,
    // aligner.hh:146-164, aligner.cc:1382-1514, not normalized; the
    // reference writes each as it goes, here they are returned, best first).
    std::vector<Alignment> suboptimal(int k, int64_t threshold = -1000000) {
        std::vector<int64_t> scores(4096);
        const int n = lcg_batch_suboptimal(batch_.get(), 0, k, threshold, scores.data(), (int32_t)scores.size());
        detail::check(n, "lcg_batch_suboptimal");
        std::vector<Alignment> out;
        const int cap = lenA_ + lenB_ + 2;
        for (int t = 0; t < n; ++t) {
            std::vector<int32_t> ea(cap), eb(cap);
            std::string sA(cap, '\0'), sB(cap, '\0');
            const int ne = lcg_batch_suboptimal_alignment(batch_.get(), t, ea.data(), eb.data(), cap, &sA[0], &sB[0],
                                                          cap, nullptr);
            detail::check(ne, "lcg_batch_suboptimal_alignment");
            Alignment a;
            for (int e = 0; e < ne; ++e) a.edges_.push_back({ea[e], eb[e]});
            a.structA_ = sA.c_str();
            a.structB_ = sB.c_str();
            a.score_ = t < (int)scores.size() ? scores[t] : 0;
            out.push_back(std::move(a));
        }
        traced_ = false;  // the batch holds the last restricted run now
        alignment_.reset();
        return out;
    }
    // The traced alignment (Aligner::get_alignment, aligner.hh:101).
    const Alignment &get_alignment() {
        if (!traced_) throw failure(LCG_E_STATE, "get_alignment before trace()");
        if (!alignment_) {
            auto a = std::make_unique<Alignment>();
            const int cap = lenA_ + lenB_ + 2;
            std::vector<int32_t> ea(cap), eb(cap);
            std::string sA(cap, '\0'), sB(cap, '\0');
            const int ne = lcg_batch_alignment(batch_.get(), 0, ea.data(), eb.data(), cap, &sA[0], &sB[0], cap);
            detail::check(ne, "lcg_batch_alignment");
            for (int k = 0; k < ne; ++k) a->edges_.push_back({ea[k], eb[k]});
            a->structA_ = sA.c_str();
            a->structB_ = sB.c_str();
            a->score_ = score();
            lcg_batch *b = batch_.get();
            a->clustal_ = detail::text_of(
                [b](char *buf, int64_t cap2) { return lcg_batch_write_clustal(b, 0, 120, 0, buf, cap2); });
            alignment_ = std::move(a);
        }
        return *alignment_;
    }
    // D(am) of every arc match in the reference's arc-match order
    // (AlignerImpl::D, aligner.cc:574-610).
    std::vector<int64_t> D() const {
        int32_t nam = 0;
        detail::check(lcg_pair_info(pair_.get(), nullptr, nullptr, &nam, nullptr, nullptr), "lcg_pair_info");
        std::vector<int64_t> d(std::max(nam, 1));
        const int n = lcg_batch_D(batch_.get(), 0, d.data(), (int32_t)d.size());
        detail::check(n, "lcg_batch_D");
        d.resize(n);
        return d;
    }

private:
    int64_t score() const {
        int64_t s = 0;
        detail::check(lcg_batch_score(batch_.get(), 0, &s), "lcg_batch_score");
        return s;
    }
    AlignerParams ap_;
    detail::PairPtr pair_;
    detail::BatchPtr batch_;
    int32_t lenA_ = 0, lenB_ = 0;
    bool traced_ = false;
    std::unique_ptr<Alignment> alignment_;
};

// Partition-function alignment (AlignerP<T>, aligner_p.hh:501-606), FP64.
template <typename T = double>
class AlignerP {
public:
    explicit AlignerP(const AlignerPParams<T> &ap) : ap_(ap) {
        if (ap.no_lonely_pairs_) throw failure(LCG_E_ARG, "no_lonely_pairs: integer alignment only");
        lcg_params p = ap.c_params();
        // locarna_p scores without stacking terms (src/locarna_p.cc:310
        // reads the joint probabilities only)
        if (p.stacking) throw failure(LCG_E_ARG, "stacking: integer alignment only");
        if (p.struct_local) throw failure(LCG_E_ARG, "struct_local: integer alignment only");
        auto pb = detail::make_pair_batch(ap, p);
        pair_ = std::move(pb.first);
        batch_ = std::move(pb.second);
        detail::check(lcg_pair_info(pair_.get(), &lenA_, &lenB_, &nam_, nullptr, nullptr), "lcg_pair_info");
    }
    void set_stream(void *stream) { detail::check(lcg_batch_set_stream(batch_.get(), stream), "lcg_batch_set_stream"); }

    // Partition function Z (AlignerP::align_inside, aligner_p.icc:413-438).
    // The library computes inside and outside in one launch; align_outside
    // then only marks the outside values as requested.
    T align_inside() {
        detail::check(lcg_batch_pf_run(batch_.get(), nullptr, nullptr), "lcg_batch_pf_run");
        double z = 0;
        detail::check(lcg_batch_pf_result(batch_.get(), 0, &z, nullptr), "lcg_batch_pf_result");
        inside_ = true;
        return z;
    }
    void align_outside() { require_inside(); }
    void compute_arcmatch_probabilities() {
        require_inside();
        am_prob_.assign(std::max(nam_, 1), 0.0);
        const int n = lcg_batch_am_probs(batch_.get(), 0, am_prob_.data(), (int32_t)am_prob_.size());
        detail::check(n, "lcg_batch_am_probs");
        am_prob_.resize(n);
    }
    void compute_basematch_probabilities(bool include_am_in_bm = false) {
        require_inside();
        include_am_in_bm_ = include_am_in_bm;
        bm_prob_.assign((size_t)(lenA_ + 1) * (lenB_ + 1), 0.0);
        detail::check(lcg_batch_bm_probs(batch_.get(), 0, include_am_in_bm ? 1 : 0, bm_prob_.data(),
                                         (int32_t)bm_prob_.size()),
                      "lcg_batch_bm_probs");
    }
    // arc-match probabilities in the reference's arc-match order
    const std::vector<double> &arcmatch_probabilities() const { return am_prob_; }
    // base-match probability of (i, j), 1-based
    double basematch_probability(int i, int j) const { return bm_prob_.at((size_t)i * (lenB_ + 1) + j); }
    std::ostream &write_arcmatch_probabilities(std::ostream &out) const {
        lcg_batch *b = batch_.get();
        return out << detail::text_of(
                   [b](char *buf, int64_t cap) { return lcg_batch_write_am_probs(b, 0, buf, cap); });
    }
    std::ostream &write_basematch_probabilities(std::ostream &out) const {
        lcg_batch *b = batch_.get();
        const int inc = include_am_in_bm_ ? 1 : 0;
        return out << detail::text_of(
                   [b, inc](char *buf, int64_t cap) { return lcg_batch_write_bm_probs(b, 0, inc, buf, cap); });
    }
    // Fragment match probability (AlignerP::compute_fragment_match_prob,
    // aligner_p.hh:605-609, aligner_p.icc:1440-1479): the probability that
    // A[i..j] and B[k..l] are matched as one block.  Like the reference,
    // computes it anew: one inside + outside run with the fragment queued.
    double compute_fragment_match_prob(int i, int j, int k, int l) {
        return compute_fragment_match_probs({{i, j, k, l}})[0];
    }
    // Several fragments in one run (no counterpart in the reference).
    std::vector<double> compute_fragment_match_probs(const std::vector<std::array<int, 4>> &frags) {
        std::vector<int32_t> q;
        for (const auto &f : frags) q.insert(q.end(), f.begin(), f.end());
        const int32_t n = (int32_t)frags.size();
        detail::check(lcg_batch_pf_set_fragments(batch_.get(), 0, n, q.data()), "lcg_batch_pf_set_fragments");
        detail::check(lcg_batch_pf_run(batch_.get(), nullptr, nullptr), "lcg_batch_pf_run");
        inside_ = true;
        std::vector<double> p(std::max(n, 1));
        const int got = lcg_batch_pf_fragment_probs(batch_.get(), 0, p.data(), (int32_t)p.size());
        detail::check(got, "lcg_batch_pf_fragment_probs");
        p.resize(got);
        return p;
    }
    // inside D(am) and outside D'(am) (AlignerP::D / Dprime)
    std::pair<std::vector<double>, std::vector<double>> D_Dprime() const {
        std::vector<double> d(std::max(nam_, 1)), dp(std::max(nam_, 1));
        const int n = lcg_batch_pf_D(batch_.get(), 0, d.data(), dp.data(), (int32_t)d.size());
        detail::check(n, "lcg_batch_pf_D");
        d.resize(n);
        dp.resize(n);
        return {d, dp};
    }

private:
    void require_inside() const {
        if (!inside_) throw failure(LCG_E_STATE, "align_inside() first");
    }
    AlignerPParams<T> ap_;
    detail::PairPtr pair_;
    detail::BatchPtr batch_;
    int32_t lenA_ = 0, lenB_ = 0, nam_ = 0;
    bool inside_ = false, include_am_in_bm_ = false;
    std::vector<double> am_prob_, bm_prob_;
};

// mlocarna's all-pairs stage (src/Utils/mlocarna:3547-3640): every pair
// (a, b), b < a, of the inputs aligned in one batch; the scores are the
// matrix mlocarna reads from the pairwise `Score:` lines.
class AllPairsAligner {
public:
    AllPairsAligner(const std::vector<RnaData> &seqs, const AlignerParams &ap) : n_((int)seqs.size()) {
        lcg_params p = ap.c_params();
        std::vector<const char *> A, B;
        for (int a = 0; a < n_; ++a)
            for (int b = 0; b < a; ++b) {
                A.push_back(seqs[a].pp_text().c_str());
                B.push_back(seqs[b].pp_text().c_str());
            }
        std::vector<lcg_pair *> pairs(A.size(), nullptr);
        lcg_prepare_opts o;
        std::memset(&o, 0, sizeof o);
        o.flags = LCG_PREPARE_DEVICE_COMPACTION | LCG_PREPARE_DEVICE_BAND;
        o.device = ap.device_;
        char err[512] = {0};
        detail::check(lcg_pairs_prepare(A.data(), B.data(), (int32_t)A.size(), &p, &o, pairs.data(), nullptr, err,
                                        sizeof err),
                      "lcg_pairs_prepare", err);
        lcg_batch *batch = nullptr;
        const int rc = lcg_batch_create(pairs.data(), (int32_t)pairs.size(), ap.device_, &batch, err, sizeof err);
        for (lcg_pair *x : pairs) lcg_pair_free(x);
        detail::check(rc, "lcg_batch_create", err);
        batch_.reset(batch);
    }
    // score + traceback of every pair in one persistent launch
    void align(bool trace = true) {
        detail::check(lcg_batch_run(batch_.get(), 0, trace ? 1 : 0, nullptr), "lcg_batch_run");
    }
    // pair index of (a, b), b < a, in mlocarna's job order
    static int index(int a, int b) { return a * (a - 1) / 2 + b; }
    int64_t score(int a, int b) const {
        int64_t s = 0;
        detail::check(lcg_batch_score(batch_.get(), index(a, b), &s), "lcg_batch_score");
        return s;
    }
    std::string clustal(int a, int b) const {
        lcg_batch *bt = batch_.get();
        const int k = index(a, b);
        return detail::text_of(
            [bt, k](char *buf, int64_t cap) { return lcg_batch_write_clustal(bt, k, 120, 0, buf, cap); });
    }

private:
    int n_;
    detail::BatchPtr batch_;
};

}  // namespace locarna_amd

#endif
