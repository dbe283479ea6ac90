// Base-pair probabilities of single sequences on the device: the step
// mlocarna runs once per input sequence before any alignment
// (compute_dotplot_rnafold_pp, mlocarna:3148-3188, which calls
// locarna_rnafold_pp: locarna_rnafold_pp.cc:209-271).  That program folds
// with ViennaRNA (RnaEnsemble, rna_ensemble.cc:141-193), keeps the pairs
// with p > --min-prob (RnaData, rna_data.cc:509-547) and writes them as a
// `#PP 2.0` text (RnaData::write_pp, rna_data.cc:1242-1350).  Here the
// partition function is the McCaskill recursion over the energy model of
// fold_model.h, run by lcg_fold_seqs (fold_kernels.hip), one workgroup per
// sequence; the text is the same writer's layout.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fold_layout.h"
#include "fold_model.h"
#include "host.hpp"
#include "locarna_amd.h"

extern "C" int lcg_launch_fold(const FoldDesc *descs, const signed char *codes, const double *tabs, double *mats,
                               double *outs, double *zout, int nseq, int nmax, hipStream_t stream);

struct lcg_fold {
    lcg_fold_params p{};
    std::vector<std::string> seq;  // normalized: upper case, T -> U
    std::vector<int> W;
    std::vector<size_t> out;       // offset of sequence k's P band in `bands`
    std::vector<double> bands;     // P and P2 bands of every sequence
    std::vector<double> Z;     // scaled partition functions
    std::vector<double> logz;  // natural logarithms of the partition functions
    float device_ms = 0, wall_ms = 0;
    int launches = 0;
    // the last pp text written per sequence and its name: a caller asks for
    // the length first and then for the text
    mutable std::vector<std::string> pp_name, pp_text;
};

namespace {

void set_err(char *err, int errlen, const std::string &msg) {
    if (err && errlen > 0) {
        std::strncpy(err, msg.c_str(), errlen - 1);
        err[errlen - 1] = 0;
    }
}

// Boltzmann factors of the loop terms, laid out as fold_layout.h describes;
// the terms depend on the pair types only through weak / strong except the
// stacks and the one-base bulges (which include a stack)
std::vector<double> factor_table(int nmax) {
    using namespace locarna_amd::fold_model;
    std::vector<double> T(LCG_FOLD_TAB_HP + 2 * (size_t)(nmax + 2), 0.0);
    const int rep[2] = {2, 1};  // a strong (CG) and a weak (AU) pair type
    for (int l1 = 0; l1 <= MAX_LOOP; ++l1)
        for (int l2 = 0; l1 + l2 <= MAX_LOOP; ++l2) {
            if (l1 + l2 < 2) continue;
            for (int wo = 0; wo < 2; ++wo)
                for (int wi = 0; wi < 2; ++wi) {
                    const int to = rep[wo], ti = rep[wi];
                    const double e = (l1 == 0 || l2 == 0) ? bulge_energy(l1 + l2, to, ti, 0.0)
                                                          : interior_energy(l1, l2, to, ti);
                    T[LCG_FOLD_TAB_LOOP + (l1 * 31 + l2) * 4 + wo * 2 + wi] = bf(e);
                }
        }
    for (int to = 1; to <= 6; ++to)
        for (int ti = 1; ti <= 6; ++ti) {
            char X, Xp, Y, Yp;
            pair_bases(to, X, Xp);
            pair_bases(ti, Y, Yp);
            const double se = stack_energy(X, Y, Yp, Xp);
            T[LCG_FOLD_TAB_STACK + to * 8 + ti] = bf(se);
            T[LCG_FOLD_TAB_BULGE1 + to * 8 + ti] = bf(bulge_energy(1, to, ti, se));
        }
    for (int w = 0; w < 2; ++w) {
        const double tp = terminal_penalty(rep[w]);
        T[LCG_FOLD_TAB_EXT + w] = bf(tp);
        T[LCG_FOLD_TAB_BRANCH + w] = bf(ML_BRANCH + tp);
        T[LCG_FOLD_TAB_TP + w] = bf(tp);
    }
    T[LCG_FOLD_TAB_MLC] = bf(ML_CLOSING + ML_BRANCH);
    for (int L = 3; L <= nmax; ++L)
        for (int w = 0; w < 2; ++w) T[LCG_FOLD_TAB_HP + L * 2 + w] = bf(hairpin_energy(L, rep[w]));
    return T;
}

signed char code_of(char c) {
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'U': return 3;
    default: return 4;
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int lcg_fold_run(const char *const *seqs, int32_t count, const lcg_fold_params *params, int32_t device,
                            lcg_fold **out, char *err, int errlen) {
    if (!seqs || count < 1 || !params || !out) return LCG_E_ARG;
    *out = nullptr;
    if (!(params->min_prob >= 0 && params->min_prob < 1) ||
        (params->max_bp_span != -1 && params->max_bp_span < 1) || (params->stacking & ~1)) {
        set_err(err, errlen, "fold parameters: min_prob in [0, 1), max_bp_span -1 or >= 1, stacking 0 or 1");
        return LCG_E_ARG;
    }
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto F = std::make_unique<lcg_fold>();
    F->p = *params;
    F->seq.resize(count);
    F->W.resize(count);
    F->out.resize(count);
    F->Z.assign(count, 1.0);
    F->logz.assign(count, 0.0);
    std::vector<FoldDesc> descs(count);
    std::vector<signed char> codes;
    size_t mat = 0, outd = 0;
    int nmax = 0;
    for (int k = 0; k < count; ++k) {
        if (!seqs[k]) return LCG_E_ARG;
        std::string s = seqs[k];
        for (char &c : s) {
            if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
            if (c == 'T') c = 'U';
        }
        const int n = (int)s.size();
        if (n > LCG_FOLD_MAX_LEN) {
            set_err(err, errlen, "sequence " + std::to_string(k) + ": length " + std::to_string(n) +
                                     " exceeds the folding limit of " + std::to_string(LCG_FOLD_MAX_LEN));
            return LCG_E_CAPACITY;
        }
        const int span = params->max_bp_span < 0 ? n : params->max_bp_span;
        const int W = std::max(1, std::min(n, span));
        FoldDesc &D = descs[k];
        D.n = n;
        D.W = W;
        D.codes = (long long)codes.size();
        codes.push_back(4);
        for (char c : s) codes.push_back(code_of(c));
        codes.push_back(4);
        const size_t sz = (size_t)(n + 2) * W;
        D.mat = (long long)mat;
        D.out = (long long)outd;
        F->out[k] = outd;
        mat += 6 * sz;
        outd += 2 * sz;
        F->W[k] = W;
        F->seq[k] = std::move(s);
        nmax = std::max(nmax, n);
    }
    const std::vector<double> tabs = factor_table(nmax);

    hipError_t e;
    auto hip_fail = [&](hipError_t x) {
        set_err(err, errlen, std::string("HIP: ") + hipGetErrorString(x));
        return LCG_E_HIP;
    };
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e);
    DevBuf dd, dc, dt, dm, dout, dz;
    if ((e = hipMalloc(&dd.p, sizeof(FoldDesc) * count)) || (e = hipMalloc(&dc.p, codes.size())) ||
        (e = hipMalloc(&dt.p, tabs.size() * sizeof(double))) || (e = hipMalloc(&dm.p, mat * sizeof(double))) ||
        (e = hipMalloc(&dout.p, outd * sizeof(double))) || (e = hipMalloc(&dz.p, count * sizeof(double))))
        return hip_fail(e);
    if ((e = hipMemcpy(dc.p, codes.data(), codes.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(dt.p, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice)))
        return hip_fail(e);
    hipEvent_t ev0, ev1;
    if ((e = hipEventCreate(&ev0))) return hip_fail(e);
    if ((e = hipEventCreate(&ev1))) {
        (void)hipEventDestroy(ev0);
        return hip_fail(e);
    }
    // scale per base: a free energy g per base (kcal/mol) makes the stored
    // values exp(L g / kT) times the Boltzmann sums of L bases.  Without a
    // given g, short sequences start unscaled and longer ones at a typical
    // ensemble free energy per base; a sequence whose values leave the
    // double range is folded again with g bisected in [-2, 1] (the
    // direction from overflow or underflow), as ViennaRNA rescales with
    // its MFE-derived pf_scale.
    using locarna_amd::fold_model::kT;
    const bool fixed = params->scale_kcal_per_nt != 0;
    std::vector<double> g(count), lo(count, -2.0), hi(count, 1.0);
    for (int k = 0; k < count; ++k)
        g[k] = fixed ? params->scale_kcal_per_nt : (F->seq[k].size() <= 400 ? 0.0 : -0.35);
    std::vector<char> ok(count, 0);
    int rc = LCG_OK;
    for (int attempt = 0; attempt < 24 && rc == LCG_OK; ++attempt) {
        for (int k = 0; k < count; ++k) descs[k].log_sigma = -g[k] / kT;
        if ((e = hipMemcpy(dd.p, descs.data(), sizeof(FoldDesc) * count, hipMemcpyHostToDevice))) {
            rc = hip_fail(e);
            break;
        }
        (void)hipEventRecord(ev0, 0);
        int lr = lcg_launch_fold((const FoldDesc *)dd.p, (const signed char *)dc.p, (const double *)dt.p,
                                 (double *)dm.p, (double *)dout.p, (double *)dz.p, count, nmax, 0);
        (void)hipEventRecord(ev1, 0);
        if (lr) {
            rc = hip_fail((hipError_t)lr);
            break;
        }
        if ((e = hipEventSynchronize(ev1))) {
            rc = hip_fail(e);
            break;
        }
        float t = 0;
        (void)hipEventElapsedTime(&t, ev0, ev1);
        F->device_ms += t;
        ++F->launches;
        if ((e = hipMemcpy(F->Z.data(), dz.p, count * sizeof(double), hipMemcpyDeviceToHost))) {
            rc = hip_fail(e);
            break;
        }
        bool all = true;
        for (int k = 0; k < count; ++k) {
            const double z = F->Z[k];
            const bool over = !(z < 1e250);  // also NaN (an infinity times 0)
            const bool under = z < 1e-250;
            ok[k] = !over && !under;
            if (ok[k]) continue;
            all = false;
            if (fixed || hi[k] - lo[k] < 1e-6) {
                set_err(err, errlen, "sequence " + std::to_string(k) + " (length " +
                                         std::to_string(F->seq[k].size()) +
                                         "): the partition function leaves the range of double at this scale");
                rc = LCG_E_CAPACITY;
                break;
            }
            if (over) hi[k] = g[k];
            else lo[k] = g[k];
            g[k] = 0.5 * (lo[k] + hi[k]);
        }
        if (all || rc) break;
    }
    if (rc == LCG_OK)
        for (int k = 0; k < count; ++k)
            if (!ok[k]) {
                set_err(err, errlen, "sequence " + std::to_string(k) + ": no scale keeps the partition function in range");
                rc = LCG_E_CAPACITY;
                break;
            }
    if (rc == LCG_OK) {
        F->bands.resize(outd);
        if ((e = hipMemcpy(F->bands.data(), dout.p, outd * sizeof(double), hipMemcpyDeviceToHost))) rc = hip_fail(e);
        for (int k = 0; k < count; ++k) F->logz[k] = std::log(F->Z[k]) + (double)F->seq[k].size() * (-g[k] / kT);
    }
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (rc) return rc;
    F->wall_ms = std::chrono::duration<float, std::milli>(clk::now() - t0).count();
    *out = F.release();
    return LCG_OK;
}

extern "C" int lcg_fold_probs(const lcg_fold *f, int32_t k, double *P, double *P2, int64_t cap) {
    if (!f || k < 0 || k >= (int)f->seq.size()) return LCG_E_ARG;
    const int n = (int)f->seq[k].size(), W = f->W[k];
    const int64_t need = (int64_t)(n + 1) * (n + 1);
    if ((!P && !P2) || cap < need) return LCG_E_ARG;
    const size_t sz = (size_t)(n + 2) * W;
    const double *bp = f->bands.data() + f->out[k];
    for (int pass = 0; pass < 2; ++pass) {
        double *M = pass ? P2 : P;
        if (!M) continue;
        std::fill(M, M + need, 0.0);
        for (int i = 1; i <= n; ++i)
            for (int j = i + 4; j <= n && j - i < W; ++j) M[(size_t)i * (n + 1) + j] = bp[pass * sz + (size_t)i * W + (j - i)];
    }
    return n;
}

extern "C" int lcg_fold_log_z(const lcg_fold *f, int32_t k, double *log_z) {
    if (!f || !log_z || k < 0 || k >= (int)f->seq.size()) return LCG_E_ARG;
    *log_z = f->logz[k];
    return LCG_OK;
}

extern "C" int64_t lcg_fold_write_pp(const lcg_fold *f, int32_t k, const char *name, char *buf, int64_t cap) {
    if (!f || !name || k < 0 || k >= (int)f->seq.size()) return LCG_E_ARG;
    if (f->pp_name.size() != f->seq.size()) {
        f->pp_name.assign(f->seq.size(), std::string());
        f->pp_text.assign(f->seq.size(), std::string());
    }
    if (!f->pp_text[k].empty() && f->pp_name[k] == name) {
        const std::string &text = f->pp_text[k];
        if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
        return (int64_t)text.size();
    }
    const std::string &s = f->seq[k];
    const int n = (int)s.size(), W = f->W[k];
    const size_t sz = (size_t)(n + 2) * W;
    const double *bp = f->bands.data() + f->out[k];
    const double cut = f->p.min_prob;
    const bool stack = f->p.stacking != 0;
    // RnaData::write_pp (rna_data.cc:1242-1350): the sequence block with a
    // name column of at least 18 characters, then the pairs above the cut
    // (i ascending, j ascending; the reference's order is that of a hash
    // map), each with its stacked joint probability where that is above
    // the cut too
    const size_t namew = std::max<size_t>(18, std::strlen(name));
    std::string text = "#PP 2.0\n\n";
    text += name;
    text.append(namew - std::strlen(name), ' ');
    text += ' ';
    text += s;
    text += "\n\n#END\n\n#SECTION BASEPAIRS\n\n#BPCUT " + locarna_amd::format_prob(cut) + "\n";
    if (stack) text += "#STACK\n";
    text += "\n";
    for (int i = 1; i <= n; ++i)
        for (int j = i + 4; j <= n && j - i < W; ++j) {
            const double p = bp[(size_t)i * W + (j - i)];
            if (!(p > cut)) continue;
            text += std::to_string(i) + " " + std::to_string(j) + " " + locarna_amd::format_prob(p);
            const double p2 = bp[sz + (size_t)i * W + (j - i)];
            if (stack && p2 > cut) text += " " + locarna_amd::format_prob(p2);
            text += "\n";
        }
    text += "\n#END\n";
    if (buf && cap > (int64_t)text.size()) std::memcpy(buf, text.c_str(), text.size() + 1);
    f->pp_name[k] = name;
    f->pp_text[k] = std::move(text);
    return (int64_t)f->pp_text[k].size();
}

extern "C" int lcg_fold_stats(const lcg_fold *f, float *device_ms, float *wall_ms) {
    if (!f) return LCG_E_ARG;
    if (device_ms) *device_ms = f->device_ms;
    if (wall_ms) *wall_ms = f->wall_ms;
    return LCG_OK;
}

extern "C" void lcg_fold_free(lcg_fold *f) { delete f; }
