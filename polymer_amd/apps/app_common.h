// app_common.h -- command line and output helpers shared by the driver
// programs written against polymer.h (the edgeMap/vertexMap interface).
//
// Graph selection (one of):
//   -rmat <scale> <edges> <seed>   device rMat a/b/c = 0.57/0.19/0.19
//   -file <path>                   PBBS AdjacencyGraph / WeightedAdjacencyGraph text
// Modifiers: -s (symmetric: symmetrise the rMat graph / the file is symmetric),
//            -b (the file is the binary <path>.config / .adj / .idx triple),
//            -wseed <k> (rMat int32 weights U[1,255] from seed k).
// Common options: -rounds <T>, -r <source>, -thr <threshold>, -o <output file>.
// The output is the result array as raw little-endian binary.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// wall-clock milliseconds: a driver times its iteration loop between two device
// synchronisations and prints "time_ms <t>"
inline double wall_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct AppArgs {
    int scale = -1;
    long edges = 0;
    unsigned long long seed = 1;
    std::string file;
    bool symmetric = false;
    bool binary = false;
    long long wseed = -1;
    int rounds = 20;
    long source = 0;
    long threshold = -1;
    std::string out;
    std::string in1, in2, in3;  // extra input arrays (app specific)
};

inline AppArgs parse_args(int argc, char **argv) {
    AppArgs a;
    for (int i = 1; i < argc; i++) {
        const std::string k = argv[i];
        auto need = [&](int c) {
            if (i + c >= argc) {
                std::fprintf(stderr, "missing value for %s\n", k.c_str());
                std::exit(2);
            }
        };
        if (k == "-rmat") {
            need(3);
            a.scale = std::atoi(argv[i + 1]);
            a.edges = std::atol(argv[i + 2]);
            a.seed = std::strtoull(argv[i + 3], nullptr, 10);
            i += 3;
        } else if (k == "-file") {
            need(1);
            a.file = argv[++i];
        } else if (k == "-s") {
            a.symmetric = true;
        } else if (k == "-b") {
            a.binary = true;
        } else if (k == "-wseed") {
            need(1);
            a.wseed = std::atoll(argv[++i]);
        } else if (k == "-rounds") {
            need(1);
            a.rounds = std::atoi(argv[++i]);
        } else if (k == "-r") {
            need(1);
            a.source = std::atol(argv[++i]);
        } else if (k == "-thr") {
            need(1);
            a.threshold = std::atol(argv[++i]);
        } else if (k == "-o") {
            need(1);
            a.out = argv[++i];
        } else if (k == "-in1") {
            need(1);
            a.in1 = argv[++i];
        } else if (k == "-in2") {
            need(1);
            a.in2 = argv[++i];
        } else if (k == "-in3") {
            need(1);
            a.in3 = argv[++i];
        } else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            std::exit(2);
        }
    }
    if (a.scale < 0 && a.file.empty()) {
        std::fprintf(stderr, "give -rmat <scale> <edges> <seed> or -file <path>\n");
        std::exit(2);
    }
    return a;
}

template <class T>
void write_output(const std::string &path, const std::vector<T> &v) {
    if (path.empty()) return;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::fclose(f);
}

template <class T>
std::vector<T> read_input(const std::string &path, size_t count) {
    std::vector<T> v(count);
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", count, path.c_str());
        std::exit(1);
    }
    std::fclose(f);
    return v;
}
