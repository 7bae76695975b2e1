// records.cpp -- records mode: perft of every position in a file, weighted by
// its occurrence count, resumable (processPerftRecords, launcher.h:1578-1695).
//
// Input: one record per line, "<FEN ...> <occurrences>" (the occurrence count
// is the last number on the line; the FEN reader ignores what follows the
// en-passant field).  Output: "<input>.op", one line per finished record,
// "<input line> <perft> <perft * occurrences>", the reference's format.  A run
// skips the records that already have an output line, so an interrupted run
// resumes where it stopped.  Unlike the reference (one record per launch),
// records are evaluated `batch` at a time through pf_perft_batch, and the
// output is appended and flushed after each batch.
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/perft_gpu_amd.h"

namespace {

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r' || s[a] == '\n')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r' || s[b - 1] == '\n')) b--;
    return s.substr(a, b - a);
}

// Last whitespace-separated token as an unsigned integer.
bool last_number(const std::string& s, uint64_t* v) {
    const size_t p = s.find_last_of(" \t");
    const std::string tok = p == std::string::npos ? s : s.substr(p + 1);
    if (tok.empty() || tok.size() > 20) return false;
    uint64_t x = 0;
    for (char c : tok) {
        if (c < '0' || c > '9') return false;
        x = x * 10 + (uint64_t)(c - '0');
    }
    *v = x;
    return true;
}

}  // namespace

extern "C" pf_status pf_records(pf_ctx* ctx, const char* in_path, int depth, int batch, uint64_t* weighted_total,
                                uint64_t* records_done) {
    if (!ctx || !in_path || depth < 0 || depth > 16 || batch < 0) return PF_EINVAL;
    if (batch == 0) batch = 256;
    const std::string op_path = std::string(in_path) + ".op";

    // finished records: complete lines of the output file (a torn last line is dropped)
    std::vector<std::string> done;
    uint64_t total = 0;
    {
        std::ifstream f(op_path, std::ios::binary);
        if (f) {
            std::stringstream ss;
            ss << f.rdbuf();
            std::string text = ss.str();
            const size_t end = text.find_last_of('\n');
            const size_t keep = end == std::string::npos ? 0 : end + 1;
            if (keep != text.size()) {
                text.resize(keep);
                std::ofstream w(op_path, std::ios::binary | std::ios::trunc);
                w << text;
                if (!w) return PF_EINVAL;
            }
            std::istringstream ls(text);
            std::string line;
            while (std::getline(ls, line)) {
                line = trim(line);
                if (line.empty()) continue;
                uint64_t v = 0;
                if (!last_number(line, &v)) return PF_EINVAL;
                total += v;
                done.push_back(line);
            }
        }
    }

    std::ifstream in(in_path);
    if (!in) return PF_EINVAL;
    FILE* op = fopen(op_path.c_str(), "ab");
    if (!op) return PF_EINVAL;
    std::vector<std::string> lines;
    std::vector<pf_pos48> pos;
    std::vector<uint64_t> occ, res;
    uint64_t processed = 0;
    size_t k = 0;
    pf_status s = PF_OK;

    auto flush = [&]() -> pf_status {
        if (lines.empty()) return PF_OK;
        res.assign(lines.size(), 0);
        const pf_status e = pf_perft_batch(ctx, pos.data(), (int)pos.size(), depth, res.data());
        if (e != PF_OK) return e;
        for (size_t i = 0; i < lines.size(); i++) {
            const uint64_t w = res[i] * occ[i];
            fprintf(op, "%s %llu %llu\n", lines[i].c_str(), (unsigned long long)res[i], (unsigned long long)w);
            total += w;
        }
        if (fflush(op) != 0) return PF_EINVAL;
        processed += lines.size();
        lines.clear();
        pos.clear();
        occ.clear();
        return PF_OK;
    };

    std::string raw;
    while (s == PF_OK && std::getline(in, raw)) {
        const std::string line = trim(raw);
        if (line.empty()) continue;
        if (k < done.size()) {
            // already evaluated: the output line must continue this record
            if (done[k].compare(0, line.size(), line) != 0) s = PF_EINVAL;
            k++;
            continue;
        }
        pf_pos48 p;
        uint64_t o = 0;
        if (!last_number(line, &o) || pf_fen_to_pos(line.c_str(), &p) != PF_OK) {
            s = PF_EINVAL;
            break;
        }
        lines.push_back(line);
        pos.push_back(p);
        occ.push_back(o);
        if ((int)lines.size() == batch) s = flush();
    }
    if (s == PF_OK && k < done.size()) s = PF_EINVAL;  // more output lines than records: another input's file
    if (s == PF_OK) s = flush();
    fclose(op);
    if (weighted_total) *weighted_total = total;
    if (records_done) *records_done = processed;
    return s;
}
