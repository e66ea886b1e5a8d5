// Minimal FITS writer for the map products: an empty primary HDU and one
// BINTABLE extension, big-endian data, 2880-byte blocks (FITS Standard 4.0).
//
// The layout follows what the reference writes through CFITSIO
// (src/fitsmod2.f90:384-407, 503-516, 543-793, 961-1061; src/output.f90):
// TTYPEn / TFORMn = "<repcount><E|D|J|K>" / TUNITn per column, keywords in
// the order the caller gives, lower-case keywords written with the HIERARCH
// convention (src/fitsmod2.f90:211-218).  Header cards are formatted per the
// standard's fixed format; they are not byte-identical to CFITSIO's.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "madam_fits.hpp"

namespace madam {

static std::string pad80(std::string s)
{
    if (s.size() > 80) s.resize(80);
    s.resize(80, ' ');
    return s;
}

static std::string quote(const std::string &v)
{
    std::string q = "'";
    for (char ch : v) {
        q += ch;
        if (ch == '\'') q += '\'';
    }
    while (q.size() < 9) q += ' ';   // at least 8 characters inside the quotes
    q += '\'';
    return q;
}

static std::string real_str(double v)
{
    char b[40];
    snprintf(b, sizeof b, "%.15G", v);
    std::string s = b;
    if (s.find('.') == std::string::npos && s.find("INF") == std::string::npos &&
        s.find("NAN") == std::string::npos) {
        size_t e = s.find('E');
        if (e == std::string::npos) s += ".";
        else s.insert(e, ".");
    }
    return s;
}

static bool is_upper_key(const std::string &k)
{
    if (k.size() > 8) return false;
    for (char ch : k)
        if (!((ch >= 'A' && ch <= 'Z') || (ch >= '0' && ch <= '9') || ch == '-' || ch == '_')) return false;
    return true;
}

static std::string card(const FitsKey &k)
{
    if (k.kind == FitsKey::COMMENT) return pad80("COMMENT " + k.comment);
    std::string value;
    bool fixed = true;   // right-justified in columns 11-30
    switch (k.kind) {
    case FitsKey::STR:
        value = quote(k.sval);
        fixed = false;
        break;
    case FitsKey::INT:
        value = std::to_string(k.ival);
        break;
    case FitsKey::REAL:
        value = real_str(k.rval);
        break;
    case FitsKey::BOOL:
        value = k.bval ? "T" : "F";
        break;
    default:
        break;
    }
    std::string c;
    if (is_upper_key(k.key)) {
        c = k.key;
        c.resize(8, ' ');
        c += "= ";
        if (fixed && value.size() < 20) c += std::string(20 - value.size(), ' ');
        c += value;
    } else {
        c = "HIERARCH " + k.key + " = " + value;
    }
    if (!k.comment.empty()) c += " / " + k.comment;
    return pad80(c);
}

static void put_be(std::vector<unsigned char> &out, const void *src, int size)
{
    const unsigned char *p = (const unsigned char *)src;
    for (int i = size - 1; i >= 0; --i) out.push_back(p[i]);
}

static int type_size(char t)
{
    switch (t) {
    case 'E': case 'J': return 4;
    case 'D': case 'K': return 8;
    default: return 0;
    }
}

std::string fits_write_bintable(const std::string &path, const std::vector<FitsColumn> &cols, int64_t nrows,
                                const std::vector<FitsKey> &keys, const std::string &extname)
{
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return "cannot create " + path + ": " + strerror(errno);
    std::string hdr;
    auto add = [&](const std::string &c) { hdr += pad80(c); };
    auto end_header = [&]() {
        add("END");
        while (hdr.size() % 2880) hdr += ' ';
        fwrite(hdr.data(), 1, hdr.size(), f);
        hdr.clear();
    };
    // primary HDU (src/fitsmod2.f90:503-516: BITPIX 8, no axes, DATE)
    add(card(FitsKey::b("SIMPLE", true, "file does conform to FITS standard")));
    add(card(FitsKey::i("BITPIX", 8, "number of bits per data pixel")));
    add(card(FitsKey::i("NAXIS", 0, "number of data axes")));
    add(card(FitsKey::b("EXTEND", true, "FITS dataset may contain extensions")));
    {
        char d[32];
        time_t t = time(nullptr);
        struct tm g;
        gmtime_r(&t, &g);
        strftime(d, sizeof d, "%Y-%m-%dT%H:%M:%S", &g);
        add(card(FitsKey::s("DATE", d, "file creation date (YYYY-MM-DDThh:mm:ss UT)")));
    }
    end_header();
    // binary table
    int64_t rowbytes = 0;
    for (const auto &c : cols) rowbytes += (int64_t)c.repcount * type_size(c.type);
    add(card(FitsKey::s("XTENSION", "BINTABLE", "binary table extension")));
    add(card(FitsKey::i("BITPIX", 8, "8-bit bytes")));
    add(card(FitsKey::i("NAXIS", 2, "2-dimensional binary table")));
    add(card(FitsKey::i("NAXIS1", rowbytes, "width of table in bytes")));
    add(card(FitsKey::i("NAXIS2", nrows, "number of rows in table")));
    add(card(FitsKey::i("PCOUNT", 0, "size of special data area")));
    add(card(FitsKey::i("GCOUNT", 1, "one data group (required keyword)")));
    add(card(FitsKey::i("TFIELDS", (int64_t)cols.size(), "number of fields in each row")));
    for (size_t i = 0; i < cols.size(); ++i) {
        std::string n = std::to_string(i + 1);
        add(card(FitsKey::s("TTYPE" + n, cols[i].name, "label for field " + n)));
        add(card(FitsKey::s("TFORM" + n, std::to_string(cols[i].repcount) + cols[i].type, "data format of field")));
        if (!cols[i].unit.empty()) add(card(FitsKey::s("TUNIT" + n, cols[i].unit, "physical unit of field")));
    }
    add(card(FitsKey::s("EXTNAME", extname, "name of this binary table extension")));
    for (const auto &k : keys) add(card(k));
    end_header();
    // rows: for each row, each column's repcount elements, big-endian
    std::vector<unsigned char> buf;
    buf.reserve((size_t)std::min<int64_t>(rowbytes * nrows, 1 << 24));
    int64_t written = 0;
    for (int64_t r = 0; r < nrows; ++r) {
        for (const auto &c : cols) {
            int sz = type_size(c.type);
            const unsigned char *base = (const unsigned char *)c.data + (size_t)r * c.repcount * sz;
            for (int e = 0; e < c.repcount; ++e) put_be(buf, base + (size_t)e * sz, sz);
        }
        if (buf.size() >= (1u << 24)) {
            fwrite(buf.data(), 1, buf.size(), f);
            written += (int64_t)buf.size();
            buf.clear();
        }
    }
    written += (int64_t)buf.size();
    int64_t padn = (2880 - written % 2880) % 2880;
    buf.insert(buf.end(), (size_t)padn, 0);
    fwrite(buf.data(), 1, buf.size(), f);
    bool bad = ferror(f) != 0;
    if (fclose(f) != 0 || bad) return "write error on " + path;
    return "";
}

}  // namespace madam
