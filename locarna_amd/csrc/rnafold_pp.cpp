// locarna_rnafold_pp on the GPU: the program mlocarna runs once per input
// sequence to write its pp file (compute_dotplot_rnafold_pp,
// src/Utils/mlocarna:3148-3188), with the reference program's options
// (src/locarna_rnafold_pp.cc:66-104) and input forms (FASTA, else a
// "name sequence" line as mlocarna pipes it, :142-187).  The folding is
// lcg_fold_run (McCaskill over this library's energy model), the text
// RnaData::write_pp's layout.  Options the model has no counterpart for are
// refused: --noLP, --in-loop, -C (structure constraints) and input with
// more than one sequence (alignment folding).  --dangling is accepted and has
// no effect (the model has no dangling-end terms).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "locarna_amd.h"

namespace {

int usage(const char *msg) {
    std::cerr << "ERROR --- " << msg << "\n"
              << "usage: locarna_rnafold_pp [-p prob] [--maxBPspan span] [--stacking] [--dangling d] [-o file] [input]\n";
    return -1;
}

// (name, sequence) records of FASTA text, else of "name sequence" lines;
// lines starting with '#' (anchor or structure annotations) are skipped
std::vector<std::pair<std::string, std::string>> read_records(const std::string &text) {
    std::vector<std::pair<std::string, std::string>> recs;
    std::istringstream in(text);
    std::string line;
    bool fasta = text.find('>') != std::string::npos;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        if (fasta) {
            if (line[0] == '>') {
                std::istringstream ls(line.substr(1));
                std::string name;
                ls >> name;
                recs.push_back({name, ""});
            } else if (!recs.empty()) {
                for (char c : line)
                    if (!std::isspace((unsigned char)c)) recs.back().second += c;
            }
        } else {
            std::istringstream ls(line);
            std::string name, seq;
            ls >> name >> seq;
            if (name.empty()) continue;
            bool found = false;
            for (auto &r : recs)
                if (r.first == name) r.second += seq, found = true;
            if (!found) recs.push_back({name, seq});
        }
    }
    return recs;
}

}  // namespace

int main(int argc, char **argv) {
    lcg_fold_params fp;
    fp.max_bp_span = -1;
    fp.stacking = 0;
    fp.min_prob = 0.001;
    fp.scale_kcal_per_nt = 0;
    std::string input = "-", output;
    for (int a = 1; a < argc; ++a) {
        const std::string o = argv[a];
        auto value = [&](const char *name) -> const char * {
            if (a + 1 >= argc) {
                usage((std::string("missing value of ") + name).c_str());
                std::exit(255);
            }
            return argv[++a];
        };
        if (o == "-p" || o == "--min-prob") fp.min_prob = std::atof(value("--min-prob"));
        else if (o.compare(0, 11, "--min-prob=") == 0) fp.min_prob = std::atof(o.c_str() + 11);
        else if (o == "--maxBPspan") fp.max_bp_span = std::atoi(value("--maxBPspan"));
        else if (o.compare(0, 12, "--maxBPspan=") == 0) fp.max_bp_span = std::atoi(o.c_str() + 12);
        else if (o == "--stacking") fp.stacking = 1;
        else if (o == "--dangling") (void)value("--dangling");
        else if (o.compare(0, 11, "--dangling=") == 0) {}
        else if (o == "-o" || o == "--output") output = value("--output");
        else if (o.compare(0, 9, "--output=") == 0) output = o.substr(9);
        else if (o == "--noLP" || o == "--in-loop" || o == "-C" || o == "--use-struct-constraints" ||
                 o == "--force-alifold")
            return usage(("option " + o + " is not supported by the GPU folding").c_str());
        else if (o == "-h" || o == "--help") {
            std::cout << "locarna_rnafold_pp -- compute RNA pair probabilities on the GPU and write in pp-format\n";
            return 0;
        } else if (!o.empty() && o[0] == '-' && o != "-")
            return usage(("unknown option " + o).c_str());
        else input = o;
    }
    if (fp.max_bp_span == 0) fp.max_bp_span = -1;
    std::string text;
    if (input == "-") {
        std::stringstream ss;
        ss << std::cin.rdbuf();
        text = ss.str();
    } else {
        std::ifstream f(input);
        if (!f) return usage(("cannot read " + input).c_str());
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    }
    const auto recs = read_records(text);
    if (recs.empty()) {
        std::cerr << "Error in input format" << std::endl;
        return -1;
    }
    if (recs.size() > 1) return usage("alignment folding (more than one sequence) is not supported");
    const char *seq = recs[0].second.c_str();
    lcg_fold *f = nullptr;
    char err[512] = {0};
    if (int rc = lcg_fold_run(&seq, 1, &fp, 0, &f, err, sizeof err)) {
        std::cerr << "ERROR --- lcg_fold_run (" << rc << "): " << err << std::endl;
        return -1;
    }
    const int64_t len = lcg_fold_write_pp(f, 0, recs[0].first.c_str(), nullptr, 0);
    std::string pp((size_t)len + 1, 0);
    lcg_fold_write_pp(f, 0, recs[0].first.c_str(), &pp[0], len + 1);
    pp.resize((size_t)len);
    lcg_fold_free(f);
    if (output.empty()) {
        std::cout << pp;
    } else {
        std::ofstream of(output);
        of << pp;
        if (!of) return usage(("cannot write " + output).c_str());
    }
    return 0;
}
