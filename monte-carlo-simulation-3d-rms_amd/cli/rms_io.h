// rms_io.h -- the reference's text file contract.
//
// Inputs as read by hpc_code/lib/main.cu:414-482 and main_PGSE.cu:440-506;
// outputs as written by main.cu:549-615 and main_PGSE.cu:595-669.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace rmsio {

struct SimParams {
    double dt = 0.0;
    int64_t TN = 0;
    int64_t NPar = 0;
    int32_t Nbvec = 0;  // narrow-pulse variant only
    int32_t Nc = 0;
    double res = 0.0;
    int32_t NPix1 = 0, NPix2 = 0, NPix3 = 0;
};

// Whitespace-separated numeric tokens of a text file.
bool read_tokens(const std::string &path, std::vector<std::string> &tok, std::string &err);

// simParamInput.txt: 9 fields (main.cu:414-419) or 8 without Nbvec
// (main_PGSE.cu:440-445).
bool read_sim_params(const std::string &path, bool pgse, SimParams &p, std::string &err);
bool read_doubles(const std::string &path, size_t n, std::vector<double> &v, std::string &err);
// fiber.txt: NPix1*NPix2*NPix3 integer labels, x-major / z fastest
// (setupRMS.m:17-23, main.cu:460-466).
bool read_labels(const std::string &path, size_t n, std::vector<uint16_t> &lab, uint32_t &maxlab, std::string &err);
// fiber.bin (no reference counterpart): the same labels as raw little-endian
// uint8 (file size n) or uint16 (file size 2n), same order.  Text is
// impractical above ~1e8 voxels.
bool read_labels_bin(const std::string &path, size_t n, std::vector<uint16_t> &lab, uint32_t &maxlab,
                     std::string &err);
bool read_int(const std::string &path, int64_t &v, std::string &err);

// Output writers (C++ stream formatting of the reference: precision 15 for
// sig0/dx2/dx4/sigRe/sig, default precision 6 for NPar_count, sim_para and
// diff_time; rows end with "\n", columns are tab-separated).
bool write_matrix(const std::string &path, const double *a, int64_t rows, int64_t cols, int precision,
                  std::string &err);
bool write_sim_para(const std::string &path, double dt, int64_t TN, int64_t NPar, double res, std::string &err);
bool write_column(const std::string &path, const std::vector<double> &v, int precision, std::string &err);

}  // namespace rmsio
