/*
 * Reader of NetCDF classic files (CDF-1, CDF-2 64-bit offset and CDF-5
 * 64-bit data), the container of the TAU mesh and solution files.  Written
 * from Unidata's published format specification ("The NetCDF Classic
 * Format Specification"): a header (magic, record count, dimension list,
 * global attributes, variable list with type, size and file offset of each
 * variable) followed by big-endian data, non-record variables contiguous at
 * their offset, record variables interleaved record by record.
 *
 * Host code only; reads variables whole, converting to the requested
 * arithmetic type.
 */
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace fj {
namespace nc {

enum Type { BYTE = 1, CHAR = 2, SHORT = 3, INT = 4, FLOAT = 5, DOUBLE = 6, UBYTE = 7, USHORT = 8, UINT = 9,
            INT64 = 10, UINT64 = 11 };

struct Error {
    std::string msg;
};

inline int type_size(int t)
{
    switch (t) {
    case BYTE: case CHAR: case UBYTE: return 1;
    case SHORT: case USHORT: return 2;
    case INT: case FLOAT: case UINT: return 4;
    case DOUBLE: case INT64: case UINT64: return 8;
    default: return 0;
    }
}

struct Dim {
    std::string name;
    uint64_t len;  /* 0 for the record dimension */
};

struct Var {
    std::string name;
    std::vector<uint64_t> dimids;
    int type = 0;
    uint64_t vsize = 0, begin = 0;
    bool record = false;
    uint64_t count = 0; /* values per record (record variables) or in total */
};

class File {
public:
    explicit File(const std::string& path) : path_(path)
    {
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw Error{"cannot open " + path};
        try {
            read_header();
        } catch (...) {
            std::fclose(f_);
            throw;
        }
    }
    ~File()
    {
        if (f_) std::fclose(f_);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;

    int version() const { return version_; }
    bool has_dim(const std::string& n) const { return find_dim(n) >= 0; }
    uint64_t dim_len(const std::string& n) const
    {
        const int d = find_dim(n);
        if (d < 0) throw Error{path_ + ": no dimension \"" + n + "\""};
        return dims_[size_t(d)].len == 0 ? numrecs_ : dims_[size_t(d)].len;
    }
    bool has_var(const std::string& n) const { return find_var(n) >= 0; }
    const Var& var(const std::string& n) const
    {
        const int v = find_var(n);
        if (v < 0) throw Error{path_ + ": no variable \"" + n + "\""};
        return vars_[size_t(v)];
    }
    /* number of values of a variable (all records) */
    uint64_t size(const std::string& n) const
    {
        const Var& v = var(n);
        return v.record ? v.count * numrecs_ : v.count;
    }
    /* shape of a variable (record dimension resolved to numrecs) */
    std::vector<uint64_t> shape(const std::string& n) const
    {
        std::vector<uint64_t> s;
        for (uint64_t d : var(n).dimids) s.push_back(dims_[size_t(d)].len == 0 ? numrecs_ : dims_[size_t(d)].len);
        return s;
    }

    /* whole variable, converted to T */
    template <class T>
    std::vector<T> get(const std::string& n)
    {
        const Var& v = var(n);
        const int ts = type_size(v.type);
        const uint64_t nrec = v.record ? numrecs_ : 1;
        std::vector<T> out(size_t(v.count * nrec));
        std::vector<unsigned char> raw(size_t(v.count) * size_t(ts));
        for (uint64_t r = 0; r < nrec; ++r) {
            const uint64_t off = v.begin + (v.record ? r * recsize_ : 0);
            if (std::fseek(f_, long(off), SEEK_SET) != 0) throw Error{path_ + ": seek failed"};
            if (!raw.empty() && std::fread(raw.data(), 1, raw.size(), f_) != raw.size())
                throw Error{path_ + ": variable \"" + n + "\" runs past the end of the file"};
            for (uint64_t i = 0; i < v.count; ++i) out[size_t(r * v.count + i)] = convert<T>(raw.data() + i * ts, v.type);
        }
        return out;
    }

private:
    std::string path_;
    FILE* f_ = nullptr;
    int version_ = 0;
    uint64_t numrecs_ = 0, recsize_ = 0;
    std::vector<Dim> dims_;
    std::vector<Var> vars_;

    int find_dim(const std::string& n) const
    {
        for (size_t i = 0; i < dims_.size(); ++i)
            if (dims_[i].name == n) return int(i);
        return -1;
    }
    int find_var(const std::string& n) const
    {
        for (size_t i = 0; i < vars_.size(); ++i)
            if (vars_[i].name == n) return int(i);
        return -1;
    }

    void bytes(void* dst, size_t n)
    {
        if (std::fread(dst, 1, n, f_) != n) throw Error{path_ + ": truncated header"};
    }
    uint32_t u32()
    {
        unsigned char b[4];
        bytes(b, 4);
        return (uint32_t(b[0]) << 24) | (uint32_t(b[1]) << 16) | (uint32_t(b[2]) << 8) | uint32_t(b[3]);
    }
    uint64_t u64()
    {
        const uint64_t hi = u32();
        return (hi << 32) | u32();
    }
    /* NON_NEG: 32-bit in CDF-1/2, 64-bit in CDF-5 */
    uint64_t non_neg() { return version_ == 5 ? u64() : u32(); }
    /* OFFSET: 32-bit in CDF-1, 64-bit otherwise */
    uint64_t offset() { return version_ == 1 ? u32() : u64(); }
    std::string name()
    {
        const uint64_t n = non_neg();
        if (n > (1u << 20)) throw Error{path_ + ": bad name length"};
        std::string s(size_t(n), '\0');
        if (n) bytes(&s[0], size_t(n));
        skip_pad(n);
        return s;
    }
    void skip_pad(uint64_t n)
    {
        const size_t pad = size_t((4 - n % 4) % 4);
        unsigned char b[4];
        if (pad) bytes(b, pad);
    }
    void skip_atts()
    {
        const uint32_t tag = u32();
        const uint64_t n = non_neg();
        if (tag == 0 && n == 0) return;
        if (tag != 0x0C) throw Error{path_ + ": bad attribute list"};
        for (uint64_t i = 0; i < n; ++i) {
            (void)name();
            const int t = int(u32());
            const uint64_t nel = non_neg();
            const int ts = type_size(t);
            if (ts == 0) throw Error{path_ + ": bad attribute type"};
            const uint64_t len = nel * uint64_t(ts);
            std::vector<unsigned char> skip(static_cast<size_t>(len));
            if (len) bytes(skip.data(), size_t(len));
            skip_pad(len);
        }
    }

    void read_header()
    {
        unsigned char magic[4];
        bytes(magic, 4);
        if (magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F') throw Error{path_ + ": not a NetCDF classic file"};
        version_ = magic[3];
        if (version_ != 1 && version_ != 2 && version_ != 5) throw Error{path_ + ": unsupported NetCDF version"};
        numrecs_ = non_neg();
        if ((version_ != 5 && numrecs_ == 0xFFFFFFFFull) || numrecs_ == ~0ull)
            throw Error{path_ + ": streaming record count not supported"};
        /* dimensions */
        uint32_t tag = u32();
        uint64_t n = non_neg();
        if (!(tag == 0 && n == 0)) {
            if (tag != 0x0A) throw Error{path_ + ": bad dimension list"};
            for (uint64_t i = 0; i < n; ++i) {
                Dim d;
                d.name = name();
                d.len = non_neg();
                dims_.push_back(d);
            }
        }
        skip_atts(); /* global attributes */
        tag = u32();
        n = non_neg();
        if (!(tag == 0 && n == 0)) {
            if (tag != 0x0B) throw Error{path_ + ": bad variable list"};
            for (uint64_t i = 0; i < n; ++i) {
                Var v;
                v.name = name();
                const uint64_t nd = non_neg();
                for (uint64_t k = 0; k < nd; ++k) {
                    const uint64_t id = non_neg();
                    if (id >= dims_.size()) throw Error{path_ + ": bad dimension id"};
                    v.dimids.push_back(id);
                }
                skip_atts();
                v.type = int(u32());
                if (type_size(v.type) == 0) throw Error{path_ + ": bad variable type"};
                v.vsize = non_neg();
                v.begin = offset();
                v.record = !v.dimids.empty() && dims_[size_t(v.dimids[0])].len == 0;
                v.count = 1;
                for (size_t k = v.record ? 1 : 0; k < v.dimids.size(); ++k) v.count *= dims_[size_t(v.dimids[k])].len;
                vars_.push_back(v);
            }
        }
        /* record size: the record variables' sizes, each padded to 4 bytes,
         * except one lone record variable, which is not padded */
        size_t nrecvars = 0;
        for (const Var& v : vars_)
            if (v.record) {
                ++nrecvars;
                recsize_ += v.vsize;
            }
        if (nrecvars == 1)
            for (const Var& v : vars_)
                if (v.record) recsize_ = v.count * uint64_t(type_size(v.type));
    }

    static uint64_t be(const unsigned char* p, int n)
    {
        uint64_t r = 0;
        for (int i = 0; i < n; ++i) r = (r << 8) | p[i];
        return r;
    }

    template <class T>
    static T convert(const unsigned char* p, int t)
    {
        switch (t) {
        case BYTE: return T(int8_t(p[0]));
        case CHAR: case UBYTE: return T(p[0]);
        case SHORT: return T(int16_t(be(p, 2)));
        case USHORT: return T(uint16_t(be(p, 2)));
        case INT: return T(int32_t(uint32_t(be(p, 4))));
        case UINT: return T(uint32_t(be(p, 4)));
        case INT64: return T(int64_t(be(p, 8)));
        case UINT64: return T(be(p, 8));
        case FLOAT: {
            const uint32_t u = uint32_t(be(p, 4));
            float f;
            std::memcpy(&f, &u, 4);
            return T(f);
        }
        case DOUBLE: {
            const uint64_t u = be(p, 8);
            double d;
            std::memcpy(&d, &u, 8);
            return T(d);
        }
        default: return T(0);
        }
    }
};

}  // namespace nc
}  // namespace fj
