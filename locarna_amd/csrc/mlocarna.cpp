// mlocarna's default pipeline on one GPU, from sequences to the multiple
// alignment: the base-pair probabilities of every input
// (compute_all_dotplots, src/Utils/mlocarna:3099-3110 -> locarna_rnafold_pp),
// the all-pairs stage (compute_pairwise_alignments, :3547-3640: one locarna
// call per pair a > b, the score read from each alignment's CLUSTAL header),
// the UPGMA guide tree of the score matrix (:2352-2420, Tree.pm:164-315) and
// the progressive alignment along it (perform_progressive_steps,
// :3660-3707).  Each stage is one entry point of this library: lcg_fold_run,
// lcg_pairs_prepare + one batch, lcg_guide_tree_upgma, lcg_progressive_run.
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "locarna_amd.h"

struct lcg_mlocarna {
    int n = 0;
    std::vector<std::string> input_pp;
    std::vector<int64_t> scores;  // n x n, row major, symmetric, 0 on the diagonal
    std::vector<int32_t> merges;
    lcg_progressive *prog = nullptr;
    float fold_ms = 0, pairs_ms = 0, progressive_ms = 0, wall_ms = 0;
    ~lcg_mlocarna() {
        if (prog) lcg_progressive_free(prog);
    }
};

namespace {

void set_err(char *err, int errlen, const std::string &msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
}

struct Pairs {
    std::vector<lcg_pair *> p;
    ~Pairs() {
        for (lcg_pair *x : p)
            if (x) lcg_pair_free(x);
    }
};

}  // namespace

extern "C" int lcg_mlocarna_run(const char *const *names, const char *const *seqs, int32_t n,
                                const lcg_fold_params *fold, const lcg_params *params, int32_t device,
                                lcg_mlocarna **out, char *err, int errlen) {
    if (!names || !seqs || n < 2 || !fold || !params || !out) return LCG_E_ARG;
    *out = nullptr;
    for (int k = 0; k < n; ++k)
        if (!names[k] || !seqs[k] || !names[k][0]) return LCG_E_ARG;
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a) { return std::chrono::duration<float, std::milli>(clk::now() - a).count(); };
    const auto t0 = clk::now();
    auto R = std::make_unique<lcg_mlocarna>();
    R->n = n;

    // base-pair probabilities of every input, written as pp texts
    lcg_fold *f = nullptr;
    if (int rc = lcg_fold_run(seqs, n, fold, device, &f, err, errlen)) return rc;
    R->input_pp.resize(n);
    for (int k = 0; k < n; ++k) {
        const int64_t len = lcg_fold_write_pp(f, k, names[k], nullptr, 0);
        std::string s((size_t)len + 1, 0);
        lcg_fold_write_pp(f, k, names[k], &s[0], len + 1);
        s.resize((size_t)len);
        R->input_pp[k] = std::move(s);
    }
    lcg_fold_free(f);
    R->fold_ms = ms(t0);

    // all pairs in mlocarna's job order (a = 0..n-1, b = 0..a-1), one batch
    const auto t1 = clk::now();
    std::vector<const char *> A, B;
    std::vector<std::pair<int, int>> ab;
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < a; ++b) {
            A.push_back(R->input_pp[a].c_str());
            B.push_back(R->input_pp[b].c_str());
            ab.push_back({a, b});
        }
    const int np = (int)ab.size();
    lcg_prepare_opts o;
    std::memset(&o, 0, sizeof o);
    o.flags = LCG_PREPARE_DEVICE_BAND;
    // the device compaction builds plain arc-match lists only
    if (!params->no_lonely_pairs && !params->stacking && !params->struct_local)
        o.flags |= LCG_PREPARE_DEVICE_COMPACTION;
    o.device = device;
    Pairs P;
    P.p.assign(np, nullptr);
    if (int rc = lcg_pairs_prepare(A.data(), B.data(), np, params, &o, P.p.data(), nullptr, err, errlen)) return rc;
    lcg_batch *batch = nullptr;
    if (int rc = lcg_batch_create(P.p.data(), np, device, &batch, err, errlen)) return rc;
    float dev = 0;
    int rc = lcg_batch_run(batch, 0, 0, &dev);
    R->scores.assign((size_t)n * n, 0);
    for (int k = 0; k < np && rc == LCG_OK; ++k) {
        int64_t sc = 0;
        rc = lcg_batch_score(batch, k, &sc);
        R->scores[(size_t)ab[k].first * n + ab[k].second] = sc;
        R->scores[(size_t)ab[k].second * n + ab[k].first] = sc;
    }
    lcg_batch_destroy(batch);
    if (rc) {
        set_err(err, errlen, "all-pairs batch failed (" + std::to_string(rc) + ")");
        return rc;
    }
    R->pairs_ms = ms(t1);

    // guide tree and progressive alignment
    const auto t2 = clk::now();
    R->merges.assign(2 * (size_t)(n - 1), 0);
    if (int rc2 = lcg_guide_tree_upgma(R->scores.data(), n, R->merges.data())) {
        set_err(err, errlen, "guide tree failed");
        return rc2;
    }
    std::vector<const char *> pp(n);
    for (int k = 0; k < n; ++k) pp[k] = R->input_pp[k].c_str();
    if (int rc3 = lcg_progressive_run(pp.data(), n, R->merges.data(), params, device, &R->prog, err, errlen))
        return rc3;
    R->progressive_ms = ms(t2);
    R->wall_ms = ms(t0);
    *out = R.release();
    return LCG_OK;
}

extern "C" int64_t lcg_mlocarna_clustal(const lcg_mlocarna *r, int32_t width, char *buf, int64_t cap) {
    if (!r || !r->prog) return LCG_E_ARG;
    return lcg_progressive_clustal(r->prog, r->n - 2, width, buf, cap);
}

extern "C" int64_t lcg_mlocarna_pp(const lcg_mlocarna *r, int32_t node, char *buf, int64_t cap) {
    if (!r || !r->prog) return LCG_E_ARG;
    return lcg_progressive_pp(r->prog, node, buf, cap);
}

extern "C" int lcg_mlocarna_scores(const lcg_mlocarna *r, int64_t *scores, int64_t cap) {
    if (!r || !scores || cap < (int64_t)r->scores.size()) return LCG_E_ARG;
    std::memcpy(scores, r->scores.data(), r->scores.size() * sizeof(int64_t));
    return r->n;
}

extern "C" int lcg_mlocarna_tree(const lcg_mlocarna *r, int32_t *merges, int32_t cap) {
    if (!r || !merges || cap < (int32_t)r->merges.size()) return LCG_E_ARG;
    std::memcpy(merges, r->merges.data(), r->merges.size() * sizeof(int32_t));
    return r->n - 1;
}

extern "C" int lcg_mlocarna_stats(const lcg_mlocarna *r, float *fold_ms, float *pairs_ms, float *progressive_ms,
                                  float *wall_ms) {
    if (!r) return LCG_E_ARG;
    if (fold_ms) *fold_ms = r->fold_ms;
    if (pairs_ms) *pairs_ms = r->pairs_ms;
    if (progressive_ms) *progressive_ms = r->progressive_ms;
    if (wall_ms) *wall_ms = r->wall_ms;
    return LCG_OK;
}

extern "C" void lcg_mlocarna_free(lcg_mlocarna *r) { delete r; }
