// Host-only FITS binary-table writer used for the map products.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace madam {

struct FitsKey {
    enum Kind { STR, INT, REAL, BOOL, COMMENT } kind = COMMENT;
    std::string key, sval, comment;
    long long ival = 0;
    double rval = 0;
    bool bval = false;
    static FitsKey s(const std::string &k, const std::string &v, const std::string &c = "")
    {
        FitsKey x;
        x.kind = STR;
        x.key = k;
        x.sval = v;
        x.comment = c;
        return x;
    }
    static FitsKey i(const std::string &k, long long v, const std::string &c = "")
    {
        FitsKey x;
        x.kind = INT;
        x.key = k;
        x.ival = v;
        x.comment = c;
        return x;
    }
    static FitsKey r(const std::string &k, double v, const std::string &c = "")
    {
        FitsKey x;
        x.kind = REAL;
        x.key = k;
        x.rval = v;
        x.comment = c;
        return x;
    }
    static FitsKey b(const std::string &k, bool v, const std::string &c = "")
    {
        FitsKey x;
        x.kind = BOOL;
        x.key = k;
        x.bval = v;
        x.comment = c;
        return x;
    }
    static FitsKey note(const std::string &c)
    {
        FitsKey x;
        x.kind = COMMENT;
        x.comment = c;
        return x;
    }
};

// One table column: nrows * repcount elements of `type` ('E' float32, 'D'
// float64, 'J' int32, 'K' int64) in host byte order, element e of row r at
// data[r * repcount + e].
struct FitsColumn {
    std::string name, unit;
    int repcount = 1;
    char type = 'E';
    const void *data = nullptr;
};

// Writes the file; returns an empty string on success, else the error.
std::string fits_write_bintable(const std::string &path, const std::vector<FitsColumn> &cols, int64_t nrows,
                                const std::vector<FitsKey> &keys, const std::string &extname);

}  // namespace madam
