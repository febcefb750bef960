// cuDQC-compatible command line (reference program.cu:7-197), one process, one GPU:
//   cudqc <graph_file> <out_gamma> <in_gamma> <min_size> <ds_sizes_file> <output_tag>
// Writes DQC-T_<tag> (raw cliques) and DQC-R_<tag> (maximal cliques, only when some
// clique was found) in the working directory and prints the reference's summary lines.
// The DS_Sizes file is read for compatibility; capacities are set by the library.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

#include "../../include/dqc.h"

int main(int argc, char* argv[]) {
    if (argc != 7) {
        std::printf("Usage: ./program1 <graph_file> <out_gamma> <in_gamma> <min_size> <ds_sizes_file> <output_file>\n");
        return 1;
    }
    {
        std::ifstream ds(argv[5]);
        if (!ds.is_open()) std::cout << "invalid data structure sizes file\n" << std::endl;
    }
    {
        std::ifstream gf(argv[1], std::ios::binary);
        if (!gf.is_open()) { std::printf("invalid graph file\n"); return 1; }
    }
    double gout = std::atof(argv[2]);
    if (gout < .5 || gout > 1) { std::printf("minimum out degree ratio must be between .5 and 1 inclusive\n"); return 1; }
    double gin = std::atof(argv[3]);
    if (gin < .5 || gin > 1) { std::printf("minimum in degree ratio must be between .5 and 1 inclusive\n"); return 1; }
    int min_size = std::atoi(argv[4]);
    if (min_size <= 1) { std::printf("minimum size must be greater than 1\n"); return 1; }
    const std::string tag = argv[6];
    const std::string raw = "DQC-T_" + tag, mx = "DQC-R_" + tag;
    auto t0 = std::chrono::steady_clock::now();
    int64_t count = 0;
    dqc_stats st{};
    dqc_status s = dqc_search_file(argv[1], gout, gin, min_size, raw.c_str(), mx.c_str(), nullptr, &count, &st);
    auto t1 = std::chrono::steady_clock::now();
    if (s != DQC_OK) { std::fprintf(stderr, "search failed with status %d\n", s); return 2; }
    long long total = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
    std::cout << "TOTAL TIME:          " << total << " ms" << std::endl << std::endl;
    std::cout << "COMPUTATION TIME:    " << (long long)st.expand_wall_ms << "ms" << std::endl;
    std::cout << "NUMBER OF CLIQUES:   " << count << std::endl;
    return 0;
}
