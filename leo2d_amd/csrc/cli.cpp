// cli.cpp -- `leo2d <deck.in> [out_prefix] [--pruned] [--device N]`
//
// One-process replacement for `mpirun -n P LEO2D deck.in` on the KPM DoS path
// (app/src/main.cpp:23-960): parses the deck, builds the geometry and job
// list, runs every job on one GPU through the C-ABI and writes <JOB_NAME>.cheb
// (and <JOB_NAME>_matrix.dat when MATRIX_SAVE > 0).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "../../include/leo2d.h"

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s deck.in [out_prefix] [--pruned] [--device N]\n", argv[0]);
        return 2;
    }
    const char* prefix = nullptr;
    int pattern = LEO2D_PATTERN_REFERENCE;
    int device = 0;
    for (int a = 2; a < argc; ++a) {
        if (!std::strcmp(argv[a], "--pruned")) pattern = LEO2D_PATTERN_PRUNED;
        else if (!std::strcmp(argv[a], "--device") && a + 1 < argc) device = std::atoi(argv[++a]);
        else prefix = argv[a];
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::printf("Input file '%s' not found, stopping, \n", argv[1]);
        return 1;
    }
    std::stringstream ss;
    ss << in.rdbuf();
    leo2d_handle* h = nullptr;
    int rc = leo2d_create(device, &h);
    if (rc != LEO2D_OK) {
        std::fprintf(stderr, "leo2d_create failed (%d)\n", rc);
        return 1;
    }
    rc = leo2d_run_deck(h, ss.str().c_str(), prefix, pattern);
    if (rc != LEO2D_OK) std::fprintf(stderr, "leo2d: error %d: %s\n", rc, leo2d_last_error(h));
    leo2d_destroy(h);
    return rc == LEO2D_OK ? 0 : 1;
}
