// rms_io.cpp -- reference text formats (see rms_io.h).
#include "rms_io.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace rmsio {

bool read_tokens(const std::string &path, std::vector<std::string> &tok, std::string &err)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    tok.clear();
    size_t i = 0, n = s.size();
    while (i < n) {
        while (i < n && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r' || s[i] == '\v' || s[i] == '\f'))
            i++;
        size_t j = i;
        while (j < n && !(s[j] == ' ' || s[j] == '\t' || s[j] == '\n' || s[j] == '\r' || s[j] == '\v' || s[j] == '\f'))
            j++;
        if (j > i) tok.emplace_back(s.substr(i, j - i));
        i = j;
    }
    return true;
}

static bool to_double(const std::string &t, double &v)
{
    char *end = nullptr;
    errno = 0;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0' && errno == 0;
}

// Integers are read the way ifstream >> int does for the files setupRMS.m
// writes ("%u"); a token that is not an integer is an error.
static bool to_int(const std::string &t, int64_t &v)
{
    char *end = nullptr;
    errno = 0;
    long long x = std::strtoll(t.c_str(), &end, 10);
    if (!end || *end != '\0' || errno != 0) return false;
    v = x;
    return true;
}

bool read_sim_params(const std::string &path, bool pgse, SimParams &p, std::string &err)
{
    std::vector<std::string> t;
    if (!read_tokens(path, t, err)) return false;
    const size_t need = pgse ? 8 : 9;
    if (t.size() < need) {
        err = path + ": expected " + std::to_string(need) + " fields, found " + std::to_string(t.size());
        return false;
    }
    size_t k = 0;
    int64_t iv = 0;
    bool ok = to_double(t[k++], p.dt);
    ok = ok && to_int(t[k++], p.TN);
    ok = ok && to_int(t[k++], p.NPar);
    if (!pgse) {
        ok = ok && to_int(t[k++], iv);
        p.Nbvec = (int32_t)iv;
    }
    ok = ok && to_int(t[k++], iv);
    p.Nc = (int32_t)iv;
    ok = ok && to_double(t[k++], p.res);
    ok = ok && to_int(t[k++], iv);
    p.NPix1 = (int32_t)iv;
    ok = ok && to_int(t[k++], iv);
    p.NPix2 = (int32_t)iv;
    ok = ok && to_int(t[k++], iv);
    p.NPix3 = (int32_t)iv;
    if (!ok) {
        err = path + ": malformed field " + std::to_string(k);
        return false;
    }
    return true;
}

bool read_doubles(const std::string &path, size_t n, std::vector<double> &v, std::string &err)
{
    std::vector<std::string> t;
    if (!read_tokens(path, t, err)) return false;
    if (t.size() < n) {
        err = path + ": expected " + std::to_string(n) + " values, found " + std::to_string(t.size());
        return false;
    }
    v.resize(n);
    for (size_t i = 0; i < n; i++) {
        if (!to_double(t[i], v[i])) {
            err = path + ": malformed value '" + t[i] + "'";
            return false;
        }
    }
    return true;
}

bool read_int(const std::string &path, int64_t &v, std::string &err)
{
    std::vector<std::string> t;
    if (!read_tokens(path, t, err)) return false;
    if (t.empty() || !to_int(t[0], v)) {
        err = path + ": expected one integer";
        return false;
    }
    return true;
}

bool read_labels_bin(const std::string &path, size_t n, std::vector<uint16_t> &lab, uint32_t &maxlab,
                     std::string &err)
{
    std::ifstream f(path, std::ios::in | std::ios::binary | std::ios::ate);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    const std::streamoff size = f.tellg();
    f.seekg(0);
    int bytes = 0;
    if (size == (std::streamoff)n)
        bytes = 1;
    else if (size == (std::streamoff)(2 * n))
        bytes = 2;
    else {
        err = path + ": size " + std::to_string((long long)size) + " is neither " + std::to_string(n) +
              " (uint8) nor " + std::to_string(2 * n) + " (uint16) bytes";
        return false;
    }
    std::vector<unsigned char> raw((size_t)size);
    if (!f.read(reinterpret_cast<char *>(raw.data()), size)) {
        err = "cannot read " + path;
        return false;
    }
    lab.resize(n);
    maxlab = 0;
    for (size_t k = 0; k < n; k++) {
        const uint32_t v = bytes == 1 ? raw[k] : (uint32_t)raw[2 * k] | ((uint32_t)raw[2 * k + 1] << 8);
        lab[k] = (uint16_t)v;
        if (v > maxlab) maxlab = v;
    }
    return true;
}

bool read_labels(const std::string &path, size_t n, std::vector<uint16_t> &lab, uint32_t &maxlab, std::string &err)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string s = ss.str();
    const char *c = s.c_str();
    const char *e = c + s.size();
    lab.resize(n);
    maxlab = 0;
    size_t k = 0;
    while (k < n) {
        while (c < e && (*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r')) c++;
        if (c >= e) break;
        if (*c < '0' || *c > '9') {
            err = path + ": label " + std::to_string(k) + " is not a non-negative integer";
            return false;
        }
        uint32_t v = 0;
        while (c < e && *c >= '0' && *c <= '9') {
            v = v * 10u + (uint32_t)(*c - '0');
            if (v > 65535u) {
                err = path + ": label above 65535";
                return false;
            }
            c++;
        }
        if (c < e && !(*c == ' ' || *c == '\t' || *c == '\n' || *c == '\r')) {
            err = path + ": malformed label " + std::to_string(k);
            return false;
        }
        lab[k++] = (uint16_t)v;
        if (v > maxlab) maxlab = v;
    }
    if (k < n) {
        err = path + ": expected " + std::to_string(n) + " labels, found " + std::to_string(k);
        return false;
    }
    return true;
}

bool write_matrix(const std::string &path, const double *a, int64_t rows, int64_t cols, int precision,
                  std::string &err)
{
    std::ofstream f(path);
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    if (precision > 0) f.precision(precision);
    for (int64_t i = 0; i < rows; i++) {
        for (int64_t j = 0; j < cols; j++) {
            f << a[i * cols + j];
            if (j == cols - 1)
                f << std::endl;
            else
                f << "\t";
        }
    }
    return true;
}

bool write_sim_para(const std::string &path, double dt, int64_t TN, int64_t NPar, double res, std::string &err)
{
    std::ofstream f(path);
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    f << dt << std::endl << TN << std::endl << NPar << std::endl;
    f << res << std::endl;
    return true;
}

bool write_column(const std::string &path, const std::vector<double> &v, int precision, std::string &err)
{
    return write_matrix(path, v.data(), (int64_t)v.size(), 1, precision, err);
}

}  // namespace rmsio
