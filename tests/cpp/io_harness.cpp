// Test harness for the CLI's file-contract code (rms_io.cpp), no GPU needed.
// usage: io_harness read <dir> <pgse 0|1> [bin] -> prints parsed parameters/labels
//        (bin: the labels come from fiber.bin instead of fiber.txt)
//        io_harness write <dir>             -> writes sample outputs
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../monte-carlo-simulation-3d-rms_amd/cli/rms_io.h"

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    std::string mode = argv[1], dir = argv[2];
    std::string err;
    if (mode == "read") {
        bool pgse = argc > 3 && std::string(argv[3]) == "1";
        rmsio::SimParams p;
        if (!rmsio::read_sim_params(dir + "/simParamInput.txt", pgse, p, err)) {
            std::cerr << err << std::endl;
            return 1;
        }
        std::printf("%.17g %lld %lld %d %d %.17g %d %d %d\n", p.dt, (long long)p.TN, (long long)p.NPar, p.Nbvec, p.Nc,
                    p.res, p.NPix1, p.NPix2, p.NPix3);
        std::vector<uint16_t> lab;
        uint32_t maxlab = 0;
        size_t n = (size_t)p.NPix1 * p.NPix2 * p.NPix3;
        const bool bin = argc > 4 && std::string(argv[4]) == "bin";
        if (!(bin ? rmsio::read_labels_bin(dir + "/fiber.bin", n, lab, maxlab, err)
                  : rmsio::read_labels(dir + "/fiber.txt", n, lab, maxlab, err))) {
            std::cerr << err << std::endl;
            return 1;
        }
        unsigned long long chk = 0;
        for (size_t i = 0; i < n; i++) chk = chk * 1000003ull + lab[i];
        std::printf("%u %llu\n", maxlab, chk);
        std::vector<double> d;
        if (!rmsio::read_doubles(dir + "/diffusivity.txt", p.Nc, d, err)) return 1;
        for (double x : d) std::printf("%.17g ", x);
        std::printf("\n");
        return 0;
    }
    if (mode == "write") {
        const double a[] = {1.0, 0.1, 1.0 / 3.0, 123456789.123456789, 1e-300, -2.5e-7, 1e6, 4096.0, 0.0};
        rmsio::write_matrix(dir + "/m15.txt", a, 3, 3, 15, err);
        rmsio::write_matrix(dir + "/m6.txt", a, 3, 3, 0, err);
        rmsio::write_sim_para(dir + "/sim_para.txt", 2e-4, 100000, 1000000, 0.1, err);
        return 0;
    }
    return 2;
}
