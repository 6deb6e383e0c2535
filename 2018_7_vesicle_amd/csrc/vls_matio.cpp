// MATLAB Level 5 MAT-file reader/writer for numeric arrays: the format of the
// reference's fixtures and run records (src/7_0_paper/2_0_Vesicle/pear.mat,
// CFL*.mat, oblate*.mat, written by MATLAB's save, e.g.
// p00_test_singlephase_relax.m:205) and of restart files.
//
// Layout (MathWorks, "MAT-File Format", Level 5): a 128-byte header (116 bytes
// of text, 8 of subsystem offset, version 0x0100, endian mark "IM"), then data
// elements.  A data element is an 8-byte tag (type, byte count) and its data
// padded to 8 bytes, or a "small" element whose type and count (<= 4 bytes)
// share the first tag word and whose data sits in the second.  miCOMPRESSED
// elements hold one zlib-compressed element.  A numeric array is an miMATRIX
// element holding array flags (class), dimensions, name and the real part,
// stored in any numeric type (MATLAB narrows integer-valued data).
//
// Reads little-endian files; every numeric class is returned as double in the
// stored (column-major) order.  Writes double arrays, optionally compressed.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/vesicle_ls.h"

namespace vls {
int api_fail(int code, const char* msg);  // vls_api.hip: sets vls_last_error()
}

namespace {

enum : uint32_t {
    miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6, miSINGLE = 7, miDOUBLE = 9,
    miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15
};
enum : int { mxDOUBLE = 6, mxUINT64 = 15 };

int fail(const std::string& m) { return vls::api_fail(VLS_ERR_INVALID, m.c_str()); }

uint32_t rd32(const uint8_t* p)
{
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

struct Elem {
    uint32_t type = 0;
    uint32_t nbytes = 0;
    const uint8_t* data = nullptr;
};

// the element at p; advances p past it (and its padding); false if malformed
bool next_elem(const uint8_t*& p, const uint8_t* end, Elem& e)
{
    if (end - p < 8) return false;
    const uint32_t w0 = rd32(p);
    if (w0 >> 16) {  // small element
        e.type = w0 & 0xffffu;
        e.nbytes = w0 >> 16;
        e.data = p + 4;
        p += 8;
        return e.nbytes <= 4;
    }
    e.type = w0;
    e.nbytes = rd32(p + 4);
    e.data = p + 8;
    const uint64_t avail = (uint64_t)(end - p - 8);
    if (e.nbytes > avail) return false;
    uint64_t adv = 8 + (uint64_t)e.nbytes;
    if (e.type != miCOMPRESSED) adv = 8 + (((uint64_t)e.nbytes + 7) & ~7ull);
    p += adv < (uint64_t)(end - p) ? adv : (uint64_t)(end - p);
    return true;
}

int type_size(uint32_t t)
{
    switch (t) {
    case miINT8: case miUINT8: return 1;
    case miINT16: case miUINT16: return 2;
    case miINT32: case miUINT32: case miSINGLE: return 4;
    case miDOUBLE: case miINT64: case miUINT64: return 8;
    default: return 0;
    }
}

template <typename S>
double conv(const uint8_t* p)
{
    S v;
    std::memcpy(&v, p, sizeof(S));
    return (double)v;
}

double to_double(uint32_t t, const uint8_t* p)
{
    switch (t) {
    case miINT8: return conv<int8_t>(p);
    case miUINT8: return conv<uint8_t>(p);
    case miINT16: return conv<int16_t>(p);
    case miUINT16: return conv<uint16_t>(p);
    case miINT32: return conv<int32_t>(p);
    case miUINT32: return conv<uint32_t>(p);
    case miSINGLE: return conv<float>(p);
    case miDOUBLE: return conv<double>(p);
    case miINT64: return conv<int64_t>(p);
    default: return conv<uint64_t>(p);
    }
}

bool inflate_all(const uint8_t* src, size_t n, std::vector<uint8_t>& out)
{
    z_stream z{};
    if (inflateInit(&z) != Z_OK) return false;
    out.assign(n * 4 + 1024, 0);
    z.next_in = const_cast<uint8_t*>(src);
    z.avail_in = (uInt)n;
    int rc = Z_OK;
    while (rc != Z_STREAM_END) {
        if (z.total_out == out.size()) out.resize(out.size() * 2);
        z.next_out = out.data() + z.total_out;
        z.avail_out = (uInt)(out.size() - z.total_out);
        rc = inflate(&z, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END && rc != Z_BUF_ERROR) break;
        if (rc == Z_BUF_ERROR && z.avail_in == 0) break;  // truncated stream
    }
    out.resize(z.total_out);
    inflateEnd(&z);
    return rc == Z_STREAM_END;
}

struct Array {
    std::string name;
    int mx_class = 0;
    bool complex = false;
    std::vector<int64_t> dims;
    Elem real;
};

// parse one miMATRIX payload; false if malformed
bool parse_matrix(const uint8_t* p, const uint8_t* end, Array& a)
{
    Elem flags, dims, name;
    if (!next_elem(p, end, flags) || flags.type != miUINT32 || flags.nbytes < 8) return false;
    const uint32_t f0 = rd32(flags.data);
    a.mx_class = (int)(f0 & 0xffu);
    a.complex = (f0 & 0x800u) != 0;
    if (a.mx_class < mxDOUBLE || a.mx_class > mxUINT64) return true;  // not numeric: listed, not readable
    if (!next_elem(p, end, dims) || dims.type != miINT32 || dims.nbytes % 4) return false;
    for (uint32_t k = 0; k < dims.nbytes / 4; ++k) a.dims.push_back((int32_t)rd32(dims.data + 4 * k));
    if (!next_elem(p, end, name) || (name.type != miINT8 && name.type != miUINT8)) return false;
    a.name.assign((const char*)name.data, name.nbytes);
    if (!next_elem(p, end, a.real) || type_size(a.real.type) == 0) return false;
    int64_t count = 1;
    for (int64_t d : a.dims) {
        if (d < 0 || (d > 0 && count > ((int64_t)1 << 40) / d)) return false;  // no element holds 2^40 values
        count *= d;
    }
    return (int64_t)a.real.nbytes == count * type_size(a.real.type);
}

struct MatFile {
    std::vector<uint8_t> bytes;
    std::vector<std::vector<uint8_t>> inflated;  // keeps decompressed payloads alive
    std::vector<Array> arrays;
};

int load(const char* path, MatFile& mf)
{
    if (!path) return fail("null path");
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fail(std::string("cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    mf.bytes.resize(sz > 0 ? (size_t)sz : 0);
    const size_t got = sz > 0 ? std::fread(mf.bytes.data(), 1, (size_t)sz, fp) : 0;
    std::fclose(fp);
    if (got != mf.bytes.size() || got < 128) return fail(std::string(path) + ": not a MAT v5 file (short)");
    const uint8_t* h = mf.bytes.data();
    if (h[126] != 'I' || h[127] != 'M') return fail(std::string(path) + ": not a little-endian MAT v5 file");
    const uint8_t* p = h + 128;
    const uint8_t* end = h + mf.bytes.size();
    mf.inflated.reserve(64);
    while (p < end) {
        Elem e;
        if (!next_elem(p, end, e)) return fail(std::string(path) + ": malformed data element");
        const uint8_t* q = e.data;
        const uint8_t* qe = e.data + e.nbytes;
        if (e.type == miCOMPRESSED) {
            mf.inflated.emplace_back();
            if (!inflate_all(e.data, e.nbytes, mf.inflated.back()))
                return fail(std::string(path) + ": bad compressed element");
            q = mf.inflated.back().data();
            qe = q + mf.inflated.back().size();
            Elem inner;
            if (!next_elem(q, qe, inner)) return fail(std::string(path) + ": malformed compressed element");
            if (inner.type != miMATRIX) continue;
            q = inner.data;
            qe = inner.data + inner.nbytes;
        } else if (e.type != miMATRIX) {
            continue;
        }
        Array a;
        if (!parse_matrix(q, qe, a)) return fail(std::string(path) + ": malformed array");
        if (a.mx_class >= mxDOUBLE && a.mx_class <= mxUINT64) mf.arrays.push_back(a);
    }
    return VLS_OK;
}

const Array* find(const MatFile& mf, const char* name)
{
    for (const Array& a : mf.arrays)
        if (!name || a.name == name) return &a;
    return nullptr;
}

void put32(std::vector<uint8_t>& b, uint32_t v)
{
    const size_t o = b.size();
    b.resize(o + 4);
    std::memcpy(b.data() + o, &v, 4);
}

void pad8(std::vector<uint8_t>& b)
{
    while (b.size() % 8) b.push_back(0);
}

}  // namespace

extern "C" {

int vls_mat_count(const char* path, int* count)
{
    if (!count) return fail("null count");
    MatFile mf;
    int rc = load(path, mf);
    if (rc) return rc;
    *count = (int)mf.arrays.size();
    return VLS_OK;
}

int vls_mat_info(const char* path, int index, char* name, int name_cap, int* ndims, int64_t* dims8, int* is_complex)
{
    if (!name || name_cap < 1 || !ndims || !dims8) return fail("null argument");
    MatFile mf;
    int rc = load(path, mf);
    if (rc) return rc;
    if (index < 0 || index >= (int)mf.arrays.size()) return fail("array index out of range");
    const Array& a = mf.arrays[index];
    if (a.dims.size() > 8) return fail("more than 8 dimensions");
    std::snprintf(name, (size_t)name_cap, "%s", a.name.c_str());
    *ndims = (int)a.dims.size();
    for (size_t k = 0; k < a.dims.size(); ++k) dims8[k] = a.dims[k];
    if (is_complex) *is_complex = a.complex ? 1 : 0;
    return VLS_OK;
}

int vls_mat_read(const char* path, const char* name, double* out, int64_t count)
{
    if (!out) return fail("null output");
    MatFile mf;
    int rc = load(path, mf);
    if (rc) return rc;
    const Array* a = find(mf, name);
    if (!a) return fail(std::string("no numeric array ") + (name ? name : "") + " in " + path);
    const int ts = type_size(a->real.type);
    const int64_t n = (int64_t)a->real.nbytes / ts;
    if (n != count) return fail("count does not match the array's " + std::to_string(n) + " elements");
    for (int64_t i = 0; i < n; ++i) out[i] = to_double(a->real.type, a->real.data + i * ts);
    return VLS_OK;
}

int vls_mat_write(const char* path, int narrays, const char* const* names, const double* const* data,
                  const int* ndims, const int64_t* dims8, int compress)
{
    if (!path || narrays < 0 || (narrays > 0 && (!names || !data || !ndims || !dims8))) return fail("null argument");
    std::vector<uint8_t> out(128, ' ');
    char text[117];
    const std::time_t now = std::time(nullptr);
    char when[64];
    std::strftime(when, sizeof when, "%a %b %d %H:%M:%S %Y", std::gmtime(&now));
    std::snprintf(text, sizeof text, "MATLAB 5.0 MAT-file, Platform: GLNXA64, Created on: %s", when);
    std::memcpy(out.data(), text, std::strlen(text));
    std::memset(out.data() + 116, 0, 8);
    out[124] = 0x00; out[125] = 0x01;  // version 0x0100, little-endian
    out[126] = 'I'; out[127] = 'M';
    for (int k = 0; k < narrays; ++k) {
        const char* nm = names[k];
        const int nd = ndims[k];
        if (!nm || !*nm || std::strlen(nm) > 63) return fail("array names are 1-63 characters");
        if (nd < 2 || nd > 8) return fail("arrays have 2 to 8 dimensions");
        int64_t count = 1;
        for (int d = 0; d < nd; ++d) {
            const int64_t v = dims8[8 * k + d];
            if (v < 0 || v > 0x7fffffff) return fail("dimension out of range");
            count *= v;
        }
        if (count > 0 && !data[k]) return fail("null data");
        if ((uint64_t)count * 8 > 0x7fffffffull) return fail("array too large for a MAT v5 element");
        std::vector<uint8_t> m;  // miMATRIX element
        put32(m, miMATRIX);
        put32(m, 0);  // size, patched below
        put32(m, miUINT32); put32(m, 8); put32(m, (uint32_t)mxDOUBLE); put32(m, 0);
        put32(m, miINT32); put32(m, (uint32_t)(4 * nd));
        for (int d = 0; d < nd; ++d) put32(m, (uint32_t)dims8[8 * k + d]);
        pad8(m);
        const uint32_t nl = (uint32_t)std::strlen(nm);
        if (nl <= 4) {
            put32(m, (nl << 16) | miINT8);
            const size_t o = m.size();
            m.resize(o + 4, 0);
            std::memcpy(m.data() + o, nm, nl);
        } else {
            put32(m, miINT8); put32(m, nl);
            m.insert(m.end(), nm, nm + nl);
            pad8(m);
        }
        put32(m, miDOUBLE); put32(m, (uint32_t)(count * 8));
        const size_t o = m.size();
        m.resize(o + (size_t)count * 8);
        if (count > 0) std::memcpy(m.data() + o, data[k], (size_t)count * 8);
        const uint32_t body = (uint32_t)(m.size() - 8);
        std::memcpy(m.data() + 4, &body, 4);
        if (compress) {
            uLongf zn = compressBound((uLong)m.size());
            std::vector<uint8_t> z(zn);
            if (compress2(z.data(), &zn, m.data(), (uLong)m.size(), 6) != Z_OK) return fail("zlib compression failed");
            put32(out, miCOMPRESSED);
            put32(out, (uint32_t)zn);
            out.insert(out.end(), z.begin(), z.begin() + (long)zn);
        } else {
            out.insert(out.end(), m.begin(), m.end());
        }
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) return fail(std::string("cannot write ") + tmp);
    const size_t put = std::fwrite(out.data(), 1, out.size(), fp);
    const bool closed = std::fclose(fp) == 0;
    const bool ok = put == out.size() && closed;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return fail(std::string("cannot write ") + path);
    }
    return VLS_OK;
}

}  // extern "C"
