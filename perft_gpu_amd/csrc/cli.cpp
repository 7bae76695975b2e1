// cli.cpp -- perft_gpu command line and UCI-style front end.
//
//   perft_gpu <fen|startpos> <depth> [backend] [tt_mib] [table_file]
//       the reference's CLI (perft.cu:29-137): runs perft for depths 1..depth
//       and prints "GPU Perft d: N, Time taken: s, nps" per depth
//       (launcher.h:1496-1498).  argv[3] selects the slider backend
//       (ks|plain|fancy|byte) instead of the reference's GPU count; with a
//       table, argv[5] names a file the table is restored from (if present)
//       and saved to afterwards.
//   perft_gpu uci
//       UCI-style loop: uci, isready, ucinewgame, position startpos|fen F
//       [moves ...], go perft N (divided perft + "Nodes searched: N"),
//       setoption name Backend value ks|plain|fancy|byte, quit.
//       The reference's UciInterface.cpp is compiled out (#if 0) and has no
//       perft verb; this is the build's own front end.
//   perft_gpu records <file> <depth> [backend] [tt_mib]
//       records mode (processPerftRecords, launcher.h:1578-1695): perft of
//       every "<FEN> <occurrences>" line, appended to <file>.op, resumable;
//       prints the weighted total.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/perft_gpu_amd.h"

static int backend_of(const std::string& s) {
    if (s == "ks") return PF_BACKEND_KS;
    if (s == "plain") return PF_BACKEND_PLAIN;
    if (s == "fancy") return PF_BACKEND_FANCY;
    if (s == "byte") return PF_BACKEND_BYTE;
    return -1;
}

static const char* kStart = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1";

static int run_cli(int argc, char** argv) {
    std::string fen = argv[1];
    if (fen == "startpos") fen = kStart;
    const int depth = atoi(argv[2]);
    pf_config cfg{};
    cfg.backend = PF_BACKEND_KS;
    if (argc > 3) {
        cfg.backend = backend_of(argv[3]);
        if (cfg.backend < 0) { fprintf(stderr, "unknown backend %s\n", argv[3]); return 2; }
    }
    if (argc > 4) cfg.tt_bytes = (uint64_t)atoll(argv[4]) << 20;
    pf_pos48 pos;
    if (pf_fen_to_pos(fen.c_str(), &pos) != PF_OK) { fprintf(stderr, "invalid FEN\n"); return 2; }
    pf_ctx* ctx = nullptr;
    int s = pf_init(&cfg, &ctx);
    if (s != PF_OK) { fprintf(stderr, "pf_init: %s\n", pf_strerror(s)); return 1; }
    // argv[5]: table file, restored before the run when it exists and saved
    // after it (the reference's perfthash.dat disk hash, launcher.h:700-789)
    const char* ttfile = argc > 5 && cfg.tt_bytes ? argv[5] : nullptr;
    if (ttfile) {
        if (FILE* f = fopen(ttfile, "rb")) {
            fclose(f);
            s = pf_tt_load(ctx, ttfile);
            if (s != PF_OK) { fprintf(stderr, "table %s: %s\n", ttfile, pf_strerror(s)); pf_destroy(ctx); return 1; }
        }
    }
    for (int d = 1; d <= depth; d++) {
        uint64_t n = 0;
        const auto t0 = std::chrono::steady_clock::now();
        s = pf_perft(ctx, &pos, d, &n);
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (s != PF_OK) { fprintf(stderr, "perft(%d): %s\n", d, pf_strerror(s)); pf_destroy(ctx); return 1; }
        printf("GPU Perft %d: %llu,   Time taken: %g seconds, nps: %llu\n", d, (unsigned long long)n, el,
               (unsigned long long)(el > 0 ? n / el : 0));
        fflush(stdout);
    }
    if (ttfile && (s = pf_tt_save(ctx, ttfile)) != PF_OK) {
        fprintf(stderr, "table %s: %s\n", ttfile, pf_strerror(s));
        pf_destroy(ctx);
        return 1;
    }
    pf_destroy(ctx);
    return 0;
}

static int run_uci() {
    pf_config cfg{};
    pf_ctx* ctx = nullptr;
    pf_pos48 pos;
    pf_fen_to_pos(kStart, &pos);
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        if (cmd == "uci") {
            printf("id name perft_gpu_amd\nid author perft_gpu_amd\n");
            printf("option name Backend type combo default ks var ks var plain var fancy var byte\n");
            printf("option name TTMiB type spin default 0 min 0 max 262144\nuciok\n");
        } else if (cmd == "isready") {
            if (!ctx) {
                const int s = pf_init(&cfg, &ctx);
                if (s != PF_OK) { printf("info string pf_init failed: %s\n", pf_strerror(s)); ctx = nullptr; }
            }
            printf("readyok\n");
        } else if (cmd == "setoption") {
            std::string tok, name, value;
            while (in >> tok) {
                if (tok == "name") in >> name;
                else if (tok == "value") in >> value;
            }
            if (name == "Backend" && backend_of(value) >= 0) {
                cfg.backend = backend_of(value);
                if (ctx) pf_set_backend(ctx, cfg.backend);
            } else if (name == "TTMiB") {
                cfg.tt_bytes = (uint64_t)atoll(value.c_str()) << 20;
                if (ctx) { pf_destroy(ctx); ctx = nullptr; }
            }
        } else if (cmd == "ucinewgame") {
            pf_fen_to_pos(kStart, &pos);
        } else if (cmd == "position") {
            std::string tok;
            in >> tok;
            std::string fen;
            if (tok == "startpos") {
                fen = kStart;
                in >> tok;
            } else if (tok == "fen") {
                std::vector<std::string> parts;
                while (in >> tok && tok != "moves") parts.push_back(tok);
                for (size_t i = 0; i < parts.size(); i++) fen += (i ? " " : "") + parts[i];
            }
            if (pf_fen_to_pos(fen.c_str(), &pos) != PF_OK) { printf("info string invalid fen\n"); continue; }
            if (tok == "moves")
                while (in >> tok)
                    if (pf_apply_uci(&pos, tok.c_str()) != PF_OK) { printf("info string illegal move %s\n", tok.c_str()); break; }
        } else if (cmd == "go") {
            std::string tok;
            int depth = 0;
            while (in >> tok)
                if (tok == "perft") in >> depth;
            if (depth < 1) { printf("info string usage: go perft <depth>\n"); continue; }
            if (!ctx && pf_init(&cfg, &ctx) != PF_OK) { printf("info string no device\n"); ctx = nullptr; continue; }
            uint16_t mv[256];
            uint64_t cnt[256];
            int n = 0;
            const auto t0 = std::chrono::steady_clock::now();
            const int s = pf_divide(ctx, &pos, depth, mv, cnt, &n);
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (s != PF_OK) { printf("info string perft failed: %s\n", pf_strerror(s)); continue; }
            uint64_t total = 0;
            for (int i = 0; i < n; i++) {
                char t[8];
                pf_move_to_uci(mv[i], t);
                printf("%s: %llu\n", t, (unsigned long long)cnt[i]);
                total += cnt[i];
            }
            printf("\nNodes searched: %llu\n", (unsigned long long)total);
            printf("info string time %.3f s nps %.0f\n", el, el > 0 ? total / el : 0.0);
        } else if (cmd == "quit") {
            break;
        }
        fflush(stdout);
    }
    if (ctx) pf_destroy(ctx);
    return 0;
}

static int run_records(int argc, char** argv) {
    const int depth = atoi(argv[3]);
    pf_config cfg{};
    cfg.backend = PF_BACKEND_KS;
    if (argc > 4) {
        cfg.backend = backend_of(argv[4]);
        if (cfg.backend < 0) { fprintf(stderr, "unknown backend %s\n", argv[4]); return 2; }
    }
    if (argc > 5) cfg.tt_bytes = (uint64_t)atoll(argv[5]) << 20;
    pf_ctx* ctx = nullptr;
    int s = pf_init(&cfg, &ctx);
    if (s != PF_OK) { fprintf(stderr, "pf_init: %s\n", pf_strerror(s)); return 1; }
    uint64_t total = 0, done = 0;
    const auto t0 = std::chrono::steady_clock::now();
    s = pf_records(ctx, argv[2], depth, 0, &total, &done);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    pf_destroy(ctx);
    if (s != PF_OK) { fprintf(stderr, "records: %s\n", pf_strerror(s)); return 1; }
    printf("Records done: %llu, Total: %g seconds, weighted perft %d: %llu\n", (unsigned long long)done, el, depth,
           (unsigned long long)total);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && !strcmp(argv[1], "uci")) return run_uci();
    if (argc >= 4 && !strcmp(argv[1], "records")) return run_records(argc, argv);
    if (argc < 3) {
        fprintf(stderr,
                "usage: perft_gpu <fen|startpos> <depth> [ks|plain|fancy|byte] [tt_mib]\n"
                "       perft_gpu uci\n"
                "       perft_gpu records <file> <depth> [ks|plain|fancy|byte] [tt_mib]\n");
        return 2;
    }
    return run_cli(argc, argv);
}
