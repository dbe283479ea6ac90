// locarna_gpu: pairwise alignment of two #PP 2.0 files on the GPU through
// the C++ facade (locarna_amd.hpp), with the call sequence of the
// reference's command-line programs: src/locarna.cc:727-785 (Aligner:
// align, trace, get_alignment; "Score: N" and the --clustal text) and, with
// --pf, src/locarna_p.cc:466-527 (AlignerP: align_inside, align_outside,
// compute_*_probabilities, the probability files; src/locarna_p.cc:537-561,
// the FRAGMENT_MATCH_PROB lines).
//
//   locarna_gpu A.pp B.pp [--indel=N] [--indel-opening=N] [--struct-weight=N]
//       [--tau=N] [--min-prob=P] [--maxBPspan=N] [--min-trace-probability=P]
//       [--sequ-local] [--free-endgaps=XXXX] [--max-diff-am=N] [--max-diff-at-am=N]
//       [--penalized=P] [--normalized=L] [--noLP] [--stacking] [--new-stacking]
//       [--exp-prob=P] [--struct-local] [--exclusion=N] [--kbest=K] [--better=T] [--clustal]
//       [--write-structure]
//   locarna_gpu --pf A.pp B.pp [options] [--min-am-prob=P] [--min-bm-prob=P]
//       [--include-am-in-bm] [--write-arcmatch-probs=FILE] [--write-basematch-probs=FILE]
//       [--fragment-match-probs="i j k l; ..."]
//
// Defaults are the reference's (locarna / locarna_p).  Exit status 0 on
// success, 1 on a library failure (the message on stderr), 2 on usage.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "locarna_amd.hpp"

namespace {

bool opt(const std::string &arg, const char *name, std::string &value) {
    const std::string key = std::string("--") + name + "=";
    if (arg.compare(0, key.size(), key) != 0) return false;
    value = arg.substr(key.size());
    return true;
}

int usage() {
    std::cerr << "usage: locarna_gpu [--pf] A.pp B.pp [options]\n";
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    using namespace locarna_amd;
    bool pf = false, clustal = false, structures = false, include_am = false, local = false, slocal = false;
    std::string files[2], free_endgaps = "----", am_file, bm_file;
    int nfiles = 0;
    ScoringParams sp;
    BandParams bp;
    int max_diff_am = -1, max_diff_at_am = -1;
    double min_am_prob = 0.001, min_bm_prob = 0.001;  // src/locarna_p.cc:143-146
    bool penalized = false, normalized = false;
    long long penalty = 0, norm_L = 0;
    bool trace_given = false, nolp = false, subopt = false, exp_prob_given = false;
    int kbest = -1;
    long long better = -1000000;  // src/locarna.cc:203-206
    std::string fragments;
    for (int a = 1; a < argc; ++a) {
        const std::string arg = argv[a];
        std::string v;
        if (arg == "--pf") pf = true;
        else if (arg == "--clustal") clustal = true;
        else if (arg == "--write-structure") structures = true;
        else if (arg == "--sequ-local") local = true;
        else if (arg == "--include-am-in-bm") include_am = true;
        else if (arg == "--noLP") nolp = true;
        else if (arg == "--stacking") sp.stacking = true;
        else if (arg == "--struct-local") slocal = true;
        else if (opt(arg, "exclusion", v)) sp.exclusion = std::atoi(v.c_str());
        else if (arg == "--new-stacking") sp.new_stacking = true;
        else if (opt(arg, "fragment-match-probs", v)) fragments = v;
        else if (opt(arg, "kbest", v)) subopt = true, kbest = std::atoi(v.c_str());
        else if (opt(arg, "better", v)) subopt = true, better = std::atoll(v.c_str());
        else if (opt(arg, "indel", v)) sp.indel = std::atoi(v.c_str());
        else if (opt(arg, "indel-opening", v)) sp.indel_opening = std::atoi(v.c_str());
        else if (opt(arg, "struct-weight", v)) sp.struct_weight = std::atoi(v.c_str());
        else if (opt(arg, "tau", v)) sp.tau_factor = std::atoi(v.c_str());
        else if (opt(arg, "match", v)) sp.match = std::atoi(v.c_str());
        else if (opt(arg, "mismatch", v)) sp.mismatch = std::atoi(v.c_str());
        else if (opt(arg, "temperature-alipf", v)) sp.temperature_alipf = std::atoi(v.c_str());
        else if (opt(arg, "exp-prob", v)) sp.exp_prob = std::atof(v.c_str()), exp_prob_given = true;
        else if (opt(arg, "min-prob", v)) bp.min_prob = std::atof(v.c_str());
        else if (opt(arg, "maxBPspan", v)) bp.max_bp_span = std::atoi(v.c_str());
        else if (opt(arg, "min-trace-probability", v)) bp.min_trace_probability = std::atof(v.c_str()), trace_given = true;
        else if (opt(arg, "max-diff", v)) bp.max_diff = std::atoi(v.c_str());
        else if (opt(arg, "max-diff-am", v)) max_diff_am = std::atoi(v.c_str());
        else if (opt(arg, "max-diff-at-am", v)) max_diff_at_am = std::atoi(v.c_str());
        else if (opt(arg, "free-endgaps", v)) free_endgaps = v;
        else if (opt(arg, "penalized", v)) penalized = true, penalty = std::atoll(v.c_str());
        else if (opt(arg, "normalized", v)) normalized = true, norm_L = std::atoll(v.c_str());
        else if (opt(arg, "min-am-prob", v)) min_am_prob = std::atof(v.c_str());
        else if (opt(arg, "min-bm-prob", v)) min_bm_prob = std::atof(v.c_str());
        else if (opt(arg, "write-arcmatch-probs", v)) am_file = v;
        else if (opt(arg, "write-basematch-probs", v)) bm_file = v;
        else if (arg.size() > 1 && arg[0] == '-') return usage();
        else if (nfiles < 2) files[nfiles++] = arg;
        else return usage();
    }
    if (nfiles != 2) return usage();
    // src/locarna.cc:407-414
    if (sp.stacking && !exp_prob_given) {
        std::cerr << "WARNING: stacking turned off. Stacking requires setting a background probability "
                     "explicitely (option --exp-prob).\n";
        sp.stacking = false;
    }
    try {
        const RnaData A = RnaData::from_file(files[0]), B = RnaData::from_file(files[1]);
        if (!pf) {
            const AlignerParams ap{AlignerParams::seqA(&A), AlignerParams::seqB(&B), AlignerParams::scoring(&sp),
                                   AlignerParams::trace_controller(&bp), AlignerParams::sequ_local(local),
                                   AlignerParams::free_endgaps(free_endgaps), AlignerParams::max_diff_am(max_diff_am),
                                   AlignerParams::max_diff_at_am(max_diff_at_am),
                                   AlignerParams::no_lonely_pairs(nolp),
                                   AlignerParams::stacking(sp.stacking || sp.new_stacking),
                                   AlignerParams::struct_local(slocal)};
            Aligner aligner(ap);
            if (subopt) {
                // src/locarna.cc:740-746: the enumeration replaces the
                // single alignment; one block per alignment
                for (const Alignment &a : aligner.suboptimal(kbest, better)) {
                    std::cout << "Score: " << a.score() << "\n";
                    std::cout << "EDGES";
                    for (const auto &e : a.alignment_edges()) std::cout << " " << e.first << ":" << e.second;
                    std::cout << "\n";
                    if (structures) {
                        std::cout << "STRUCTURE_A " << a.dot_bracket_structureA() << "\n";
                        std::cout << "STRUCTURE_B " << a.dot_bracket_structureB() << "\n";
                    }
                    std::cout << "\n";
                }
                return 0;
            }
            int64_t score;
            if (normalized) score = aligner.normalized_align(norm_L);
            else if (penalized) score = aligner.penalized_align(penalty);
            else {
                score = aligner.align();
                aligner.trace();
            }
            const Alignment &alignment = aligner.get_alignment();
            if (is_neg_infty(score)) std::cout << "Score: -inf\n";
            else std::cout << "Score: " << score << "\n";
            std::cout << "EDGES";
            for (const auto &e : alignment.alignment_edges()) std::cout << " " << e.first << ":" << e.second;
            std::cout << "\n";
            if (structures) {
                std::cout << "STRUCTURE_A " << alignment.dot_bracket_structureA() << "\n";
                std::cout << "STRUCTURE_B " << alignment.dot_bracket_structureB() << "\n";
            }
            if (clustal) alignment.write_clustal(std::cout);
        } else {
            if (!trace_given) bp.min_trace_probability = 1e-5;
            const AlignerPParams<double> ap{AlignerParams::seqA(&A), AlignerParams::seqB(&B),
                                            AlignerParams::scoring(&sp), AlignerParams::trace_controller(&bp),
                                            AlignerPParams<double>::min_am_prob(min_am_prob),
                                            AlignerPParams<double>::min_bm_prob(min_bm_prob),
                                            AlignerParams::no_lonely_pairs(nolp),
                                            AlignerParams::struct_local(slocal)};
            AlignerP<double> aligner(ap);
            std::cout << "Run inside algorithm.\n";
            const double z = aligner.align_inside();
            std::cout << "Partition function: " << z << "\n";
            std::cout << "Run outside algorithm.\n";
            aligner.align_outside();
            std::cout << "Compute probabilities.\n";
            aligner.compute_arcmatch_probabilities();
            aligner.compute_basematch_probabilities(include_am);
            if (!am_file.empty()) {
                std::ofstream out(am_file);
                aligner.write_arcmatch_probabilities(out);
            }
            if (!bm_file.empty()) {
                std::ofstream out(bm_file);
                aligner.write_basematch_probabilities(out);
            }
            // fragment_list ::= empty | pos pos pos pos {; pos pos pos pos}
            std::vector<std::array<int, 4>> frags;
            {
                std::istringstream in(fragments);
                std::string item;
                while (std::getline(in, item, ';')) {
                    std::istringstream one(item);
                    std::array<int, 4> f;
                    if (!(one >> f[0] >> f[1] >> f[2] >> f[3])) {
                        if (item.find_first_not_of(" \t") == std::string::npos) continue;
                        std::cerr << "WARNING: expected 4 positions or end when parsing argument "
                                     "fragment-match-probs\"" << fragments << "\"\n";
                        break;
                    }
                    frags.push_back(f);
                }
            }
            if (!frags.empty()) {
                const std::vector<double> p = aligner.compute_fragment_match_probs(frags);
                for (size_t q = 0; q < frags.size(); ++q)
                    std::cout << "FRAGMENT_MATCH_PROB " << frags[q][0] << " " << frags[q][1] << " " << frags[q][2]
                              << " " << frags[q][3] << " : " << p[q] << "\n";
            }
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.17g", z);
            std::cout << "Z_EXACT " << buf << "\n";
        }
    } catch (failure &f) {
        std::cerr << "locarna_gpu: " << f.what() << "\n";
        return 1;
    }
    return 0;
}
