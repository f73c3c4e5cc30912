// Bakes the oracle's libbsdl energy tables into a file (run by `make`):
//   genluts <path>
// TEST INFRASTRUCTURE ONLY (see ../oracle_core.h).
#include <cstdio>
#include <thread>
#include <vector>

#include "../mtx.h"

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: genluts <output>\n");
        return 2;
    }
    std::vector<float> t(oref::mtx::kLutFloats);
    unsigned hc = std::thread::hardware_concurrency();
    oref::mtx::bake_luts(t.data(), hc ? (int)(hc < 16 ? hc : 16) : 4);
    return oref::mtx::save_luts(argv[1], t.data()) ? 0 : 1;
}
