// mlocarna's progressive stage on the GPU: the guide tree from the all-pairs
// score matrix and the alignment of alignments along it.
//
// mlocarna (src/Utils/mlocarna:2352-2420, 3660-3707) builds a UPGMA tree
// from the pairwise scores (lib/perl/MLocarna/Tree.pm:164-315), then walks
// it bottom up and runs one locarna process per inner node: the pp files of
// the two children are aligned and the consensus pp of the result (locarna
// --pp, rna_data.cc:1475-1570) is the node's input to its parent
// (call_locarna, mlocarna:3462-3500).  Here the nodes whose children are
// done run together: each round prepares their pairs on host threads,
// aligns them in one batch per kind of score table (pairs of single
// sequences read sigma by symbol code, pairs with an alignment profile on a
// side from a column-pair table) and writes their consensus pp texts.  The
// result of a node does not depend on the order of the rounds.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "host.hpp"
#include "locarna_amd.h"

struct lcg_progressive {
    int n = 0;
    std::vector<std::string> text;  // pp text of every node: leaves 0..n-1, merge t at n + t
    std::vector<int64_t> score;     // per merge
    std::vector<int> single;        // per merge: both children single sequences
    int rounds = 0, batches = 0;
    float device_ms = 0, wall_ms = 0, prepare_ms = 0;
};

namespace {

void set_err(char *err, int errlen, const std::string &msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
}

// (name, row) lines of the sequence block of a pp text
std::vector<std::pair<std::string, std::string>> pp_rows(const std::string &pp) {
    std::vector<std::pair<std::string, std::string>> rows;
    std::istringstream in(pp);
    std::string line;
    std::getline(in, line);  // #PP 2.0
    while (std::getline(in, line)) {
        if (line.compare(0, 4, "#END") == 0) break;
        if (line.empty() || line[0] == '#' || line[0] == ' ' || line[0] == '\t') continue;
        std::istringstream ls(line);
        std::string name, seq;
        ls >> name >> seq;
        bool found = false;
        for (auto &r : rows)
            if (r.first == name) r.second += seq, found = true;
        if (!found) rows.push_back({name, seq});
    }
    return rows;
}

struct PairGuard {
    std::vector<lcg_pair *> p;
    ~PairGuard() {
        for (lcg_pair *x : p)
            if (x) lcg_pair_free(x);
    }
};
struct BatchGuard {
    lcg_batch *b = nullptr;
    ~BatchGuard() {
        if (b) lcg_batch_destroy(b);
    }
};

}  // namespace

extern "C" int lcg_guide_tree_upgma(const int64_t *scores, int32_t n, int32_t *merges) {
    if (!scores || n < 1 || (n > 1 && !merges)) return LCG_E_ARG;
    // distances max - score; max over the upper triangle and the entry
    // [0][0], which mlocarna has set to 0 (mlocarna:2357-2360, Tree.pm:265-291)
    double mx = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) mx = std::max(mx, (double)scores[(size_t)i * n + j]);
    std::vector<double> dist((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) dist[(size_t)i * n + j] = mx - (i == j ? 0.0 : (double)scores[(size_t)i * n + j]);
    auto D = [&](int a, int b) -> double & { return dist[(size_t)a * n + b]; };
    // the cluster list in the order the reference keeps it: the merged
    // cluster takes slot 0 under the id of its first member, the first
    // cluster moves into the first member's slot, the last into the
    // second member's slot, and the list loses its last slot
    std::vector<int> clusters(n), size(n, 1), node(n);
    for (int i = 0; i < n; ++i) clusters[i] = node[i] = i;
    for (int t = 0; t + 1 < n; ++t) {
        const int last = n - 1 - t;
        int mi = -1, mj = -1;
        double md = 1e10;
        for (int i = 0; i <= last; ++i)
            for (int j = i + 1; j <= last; ++j)
                if (D(clusters[i], clusters[j]) < md) {
                    md = D(clusters[i], clusters[j]);
                    mi = i;
                    mj = j;
                }
        if (mi < 0) return LCG_E_ARG;  // distances beyond the reference's 1e10 sentinel
        const int ci = clusters[mi], cj = clusters[mj];
        clusters[mj] = clusters[last];
        clusters[mi] = clusters[0];
        clusters[0] = ci;
        for (int i = 1; i < last; ++i) {
            D(clusters[0], clusters[i]) = (size[ci] * D(ci, clusters[i]) + size[cj] * D(cj, clusters[i])) /
                                          (double)(size[ci] + size[cj]);
            D(clusters[i], clusters[0]) = D(clusters[0], clusters[i]);
        }
        merges[2 * t] = node[ci];
        merges[2 * t + 1] = node[cj];
        node[ci] = n + t;
        size[ci] += size[cj];
    }
    return LCG_OK;
}

extern "C" int lcg_progressive_run(const char *const *pp, int32_t n, const int32_t *merges, const lcg_params *params,
                                   int32_t device, lcg_progressive **out, char *err, int errlen) {
    if (!pp || n < 1 || !params || !out || (n > 1 && !merges)) return LCG_E_ARG;
    *out = nullptr;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto R = std::make_unique<lcg_progressive>();
    R->n = n;
    const int nm = n - 1;
    R->text.resize((size_t)n + nm);
    for (int i = 0; i < n; ++i) {
        if (!pp[i]) return LCG_E_ARG;
        R->text[i] = pp[i];
    }
    R->score.assign(nm, 0);
    R->single.assign(nm, 0);
    // a binary tree over the n leaves: every node is a child once, children
    // come before their merge
    std::vector<int> used((size_t)n + nm, 0);
    for (int t = 0; t < nm; ++t)
        for (int c = 0; c < 2; ++c) {
            const int x = merges[2 * t + c];
            if (x < 0 || x >= n + t || used[x]++) {
                set_err(err, errlen, "merges: not a binary tree over the leaves (merge " + std::to_string(t) + ")");
                return LCG_E_ARG;
            }
        }
    std::vector<char> done((size_t)n + nm, 0);
    std::fill(done.begin(), done.begin() + n, 1);
    int left = nm;
    while (left > 0) {
        std::vector<int> ready;
        for (int t = 0; t < nm; ++t)
            if (!done[n + t] && done[merges[2 * t]] && done[merges[2 * t + 1]]) ready.push_back(t);
        const int k = (int)ready.size();
        std::vector<const char *> A(k), B(k);
        for (int x = 0; x < k; ++x) {
            A[x] = R->text[merges[2 * ready[x]]].c_str();
            B[x] = R->text[merges[2 * ready[x] + 1]].c_str();
        }
        // pairs of two single sequences are prepared with the device band
        // and, where the parameters allow it, the device arc-match
        // compaction; pairs with a profile on a side with the device band
        // call, which bands them on the host
        const bool compaction = !params->no_lonely_pairs && !params->stacking && !params->struct_local;
        PairGuard P;
        P.p.assign(k, nullptr);
        const auto tp = clk::now();
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<int> idx;
            for (int x = 0; x < k; ++x) {
                const bool single = pp_rows(A[x]).size() == 1 && pp_rows(B[x]).size() == 1;
                if (single == (pass == 0)) idx.push_back(x);
            }
            if (idx.empty()) continue;
            std::vector<const char *> a, b;
            for (int x : idx) a.push_back(A[x]), b.push_back(B[x]);
            std::vector<lcg_pair *> got(idx.size(), nullptr);
            lcg_prepare_opts o;
            std::memset(&o, 0, sizeof o);
            o.flags = LCG_PREPARE_DEVICE_BAND | (pass == 0 && compaction ? LCG_PREPARE_DEVICE_COMPACTION : 0);
            o.device = device;
            const int rc = lcg_pairs_prepare(a.data(), b.data(), (int)idx.size(), params, &o, got.data(), nullptr,
                                             err, errlen);
            for (size_t y = 0; y < idx.size(); ++y) P.p[idx[y]] = got[y];
            if (rc) return rc;
        }
        R->prepare_ms += std::chrono::duration<float, std::milli>(clk::now() - tp).count();
        // one batch per kind of sigma table
        std::vector<int> kind(k);
        for (int x = 0; x < k; ++x) {
            int32_t ra = 0, rb = 0;
            if (int rc = lcg_pair_score_tables(P.p[x], nullptr, nullptr, nullptr, &ra, &rb)) return rc;
            kind[x] = (ra > 1 || rb > 1) ? 1 : 0;
            R->single[ready[x]] = kind[x] == 0;
        }
        for (int kd = 0; kd < 2; ++kd) {
            std::vector<int> idx;
            for (int x = 0; x < k; ++x)
                if (kind[x] == kd) idx.push_back(x);
            if (idx.empty()) continue;
            std::vector<lcg_pair *> bp;
            for (int x : idx) bp.push_back(P.p[x]);
            BatchGuard G;
            if (int rc = lcg_batch_create(bp.data(), (int32_t)bp.size(), device, &G.b, err, errlen)) return rc;
            float ms = 0;
            if (int rc = lcg_batch_run(G.b, 0, 1, &ms)) {
                set_err(err, errlen, "lcg_batch_run failed (" + std::to_string(rc) + ")");
                return rc;
            }
            R->device_ms += ms;
            ++R->batches;
            for (size_t y = 0; y < idx.size(); ++y) {
                const int t = ready[idx[y]];
                int64_t sc = 0;
                if (int rc = lcg_batch_score(G.b, (int32_t)y, &sc)) return rc;
                R->score[t] = sc;
                const int64_t len = lcg_batch_write_consensus_pp(G.b, (int32_t)y, nullptr, 0);
                if (len < 0) return (int)len;
                std::string s((size_t)len + 1, 0);
                lcg_batch_write_consensus_pp(G.b, (int32_t)y, &s[0], len + 1);
                s.resize((size_t)len);
                R->text[n + t] = std::move(s);
            }
        }
        for (int t : ready) done[n + t] = 1;
        left -= k;
        ++R->rounds;
    }
    R->wall_ms = std::chrono::duration<float, std::milli>(clk::now() - t0).count();
    *out = R.release();
    return LCG_OK;
}

extern "C" int64_t lcg_progressive_pp(const lcg_progressive *r, int32_t node, char *buf, int64_t cap) {
    if (!r || node < 0 || node >= (int)r->text.size()) return LCG_E_ARG;
    const std::string &s = r->text[node];
    if (buf && cap > (int64_t)s.size()) std::memcpy(buf, s.c_str(), s.size() + 1);
    return (int64_t)s.size();
}

extern "C" int64_t lcg_progressive_clustal(const lcg_progressive *r, int32_t t, int32_t width, char *buf,
                                           int64_t cap) {
    if (!r || t < 0 || t >= r->n - 1 || width < 1) return LCG_E_ARG;
    // the text locarna writes with --clustal for the merge (main_helper.icc:
    // 557-580): the score in the header for two single sequences only
    const auto rows = pp_rows(r->text[r->n + t]);
    std::vector<std::pair<std::string, const std::string *>> lines;
    for (const auto &x : rows) lines.push_back({x.first, &x.second});
    std::string text = "CLUSTAL W --- LocARNA 2.0.1";
    if (r->single[t]) text += " --- Score: " + std::to_string(r->score[t]);
    text += "\n\n" + locarna_amd::clustal_blocks(lines, (size_t)width);
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    return (int64_t)text.size();
}

extern "C" int lcg_progressive_scores(const lcg_progressive *r, int64_t *scores, int32_t cap) {
    if (!r || !scores || cap < r->n - 1) return LCG_E_ARG;
    for (int t = 0; t < r->n - 1; ++t) scores[t] = r->score[t];
    return r->n - 1;
}

extern "C" int lcg_progressive_stats(const lcg_progressive *r, int32_t *rounds, int32_t *batches, float *device_ms,
                                     float *prepare_ms, float *wall_ms) {
    if (!r) return LCG_E_ARG;
    if (rounds) *rounds = r->rounds;
    if (batches) *batches = r->batches;
    if (device_ms) *device_ms = r->device_ms;
    if (prepare_ms) *prepare_ms = r->prepare_ms;
    if (wall_ms) *wall_ms = r->wall_ms;
    return LCG_OK;
}

extern "C" void lcg_progressive_free(lcg_progressive *r) { delete r; }
