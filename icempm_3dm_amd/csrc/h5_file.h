// h5_file.h — a small reader of the HDF5 file format subset used by the
// reference's point clouds and snapshots (version 0/1 superblock, version 1
// object headers, symbol-table groups, contiguous or chunked layout version 3
// with deflate/shuffle/fletcher32 filters, little-endian fixed-point and IEEE
// types), written from the published HDF5 file format specification. Every
// read is bounds-checked against the file size; anything outside the subset
// throws H5Error, never a guess. Host code only.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <utility>
#include <vector>

namespace icempm {
namespace h5 {

struct H5Error {
    std::string msg;
};

[[noreturn]] inline void bad(const std::string &m) { throw H5Error{m}; }

struct File {
    std::vector<uint8_t> d;
    int so = 8, sl = 8; // size of offsets / lengths

    uint64_t u(uint64_t at, int n) const {
        if (n < 1 || n > 8 || at > d.size() || d.size() - at < (uint64_t)n) bad("read past the end of the file");
        uint64_t v = 0;
        for (int k = 0; k < n; k++) v |= (uint64_t)d[at + k] << (8 * k);
        return v;
    }
    uint64_t off(uint64_t at) const { return u(at, so); }
    uint64_t len(uint64_t at) const { return u(at, sl); }
    void need(uint64_t at, uint64_t n) const {
        if (at > d.size() || d.size() - at < n) bad("structure extends past the end of the file");
    }
    bool sig(uint64_t at, const char *s) const {
        need(at, 4);
        return std::memcmp(&d[at], s, 4) == 0;
    }
    std::string cstr(uint64_t at, uint64_t maxlen) const {
        std::string s;
        for (uint64_t k = 0; k < maxlen; k++) {
            need(at + k, 1);
            if (d[at + k] == 0) return s;
            s.push_back((char)d[at + k]);
        }
        bad("unterminated name");
    }
};

constexpr uint64_t kUndef = ~0ull;

struct Datatype {
    int cls = -1;  // 0 fixed-point, 1 floating-point
    int size = 0;  // bytes
    bool big_endian = false;
    bool is_signed = false;
};

struct Filter {
    int id;
    std::vector<uint32_t> cd;
};

struct Dataset {
    std::vector<uint64_t> dims;
    Datatype type;
    int layout_class = -1; // 1 contiguous, 2 chunked
    uint64_t addr = kUndef, size = 0;
    std::vector<uint32_t> chunk; // chunk dims incl. the element-size dimension
    std::vector<Filter> filters;
    struct Attr {
        std::string name;
        Datatype type;
        std::vector<uint8_t> data;
    };
    std::vector<Attr> attrs;
};

inline Datatype parse_datatype(const File &f, uint64_t at) {
    Datatype t;
    const uint64_t b0 = f.u(at, 1);
    t.cls = (int)(b0 & 0x0f);
    const uint64_t bits = f.u(at + 1, 3);
    t.size = (int)f.u(at + 4, 4);
    if (t.cls == 0 || t.cls == 1) {
        t.big_endian = (bits & 1) != 0;
        t.is_signed = t.cls == 0 && (bits & 0x08) != 0;
        if (t.cls == 1 && (bits & 0x40)) bad("VAX floating-point order is not supported");
    }
    return t;
}

inline std::vector<uint64_t> parse_dataspace(const File &f, uint64_t at) {
    const int ver = (int)f.u(at, 1);
    const int rank = (int)f.u(at + 1, 1);
    uint64_t p;
    if (ver == 1) p = at + 8;
    else if (ver == 2) p = at + 4;
    else bad("unsupported dataspace message version");
    std::vector<uint64_t> dims;
    for (int k = 0; k < rank; k++) dims.push_back(f.len(p + (uint64_t)k * f.sl));
    return dims;
}

inline void parse_filters(const File &f, uint64_t at, Dataset &ds) {
    const int ver = (int)f.u(at, 1);
    const int nf = (int)f.u(at + 1, 1);
    uint64_t p = (ver == 1) ? at + 8 : at + 2;
    if (ver != 1 && ver != 2) bad("unsupported filter pipeline version");
    for (int k = 0; k < nf; k++) {
        Filter flt;
        flt.id = (int)f.u(p, 2);
        p += 2;
        uint64_t namelen = 0;
        if (ver == 1 || flt.id >= 256) {
            namelen = f.u(p, 2);
            p += 2;
        }
        p += 2; // flags
        const int ncd = (int)f.u(p, 2);
        p += 2;
        if (ver == 1) namelen = (namelen + 7) / 8 * 8;
        p += namelen;
        for (int c = 0; c < ncd; c++) flt.cd.push_back((uint32_t)f.u(p + 4 * (uint64_t)c, 4));
        p += 4 * (uint64_t)ncd;
        if (ver == 1 && (ncd & 1)) p += 4;
        ds.filters.push_back(flt);
    }
}

inline void parse_attribute(const File &f, uint64_t at, Dataset &ds) {
    const int ver = (int)f.u(at, 1);
    const uint64_t nsz = f.u(at + 2, 2), tsz = f.u(at + 4, 2), ssz = f.u(at + 6, 2);
    uint64_t p = at + 8;
    if (ver == 3) p += 1; // name character set
    else if (ver != 1 && ver != 2) bad("unsupported attribute message version");
    auto pad = [&](uint64_t n) { return ver == 1 ? (n + 7) / 8 * 8 : n; };
    Dataset::Attr a;
    a.name = f.cstr(p, nsz);
    p += pad(nsz);
    a.type = parse_datatype(f, p);
    p += pad(tsz);
    const std::vector<uint64_t> dims = parse_dataspace(f, p);
    p += pad(ssz);
    uint64_t count = 1;
    for (uint64_t d : dims) {
        if (d != 0 && count > f.d.size() / d) bad("attribute larger than the file");
        count *= d;
    }
    if (a.type.size < 0 || (a.type.size > 0 && count > f.d.size() / (uint64_t)a.type.size))
        bad("attribute larger than the file");
    const uint64_t nbytes = count * (uint64_t)a.type.size;
    f.need(p, nbytes);
    a.data.assign(f.d.begin() + (long)p, f.d.begin() + (long)(p + nbytes));
    ds.attrs.push_back(a);
}

inline void parse_messages(const File &f, uint64_t p, uint64_t end, Dataset &ds, std::vector<std::pair<uint64_t, uint64_t>> &more) {
    while (p + 8 <= end) {
        const int type = (int)f.u(p, 2);
        const uint64_t size = f.u(p + 2, 2);
        const uint64_t body = p + 8;
        f.need(body, size);
        switch (type) {
        case 0x0001: ds.dims = parse_dataspace(f, body); break;
        case 0x0003: ds.type = parse_datatype(f, body); break;
        case 0x0008: {
            const int ver = (int)f.u(body, 1);
            if (ver != 3) bad("only data layout message version 3 is supported");
            ds.layout_class = (int)f.u(body + 1, 1);
            if (ds.layout_class == 1) {
                ds.addr = f.off(body + 2);
                ds.size = f.len(body + 2 + f.so);
            } else if (ds.layout_class == 2) {
                const int nd = (int)f.u(body + 2, 1);
                ds.addr = f.off(body + 3);
                for (int k = 0; k < nd; k++) ds.chunk.push_back((uint32_t)f.u(body + 3 + f.so + 4 * (uint64_t)k, 4));
            } else {
                bad("compact or virtual dataset layout is not supported");
            }
            break;
        }
        case 0x000B: parse_filters(f, body, ds); break;
        case 0x000C: parse_attribute(f, body, ds); break;
        case 0x0010: more.push_back({f.off(body), f.len(body + f.so)}); break;
        default: break;
        }
        p = body + size;
    }
}

inline Dataset parse_object(const File &f, uint64_t addr) {
    if (f.u(addr, 1) != 1) bad("only version 1 object headers are supported");
    const uint64_t hsize = f.u(addr + 8, 4);
    Dataset ds;
    std::vector<std::pair<uint64_t, uint64_t>> blocks{{addr + 16, hsize}};
    for (size_t b = 0; b < blocks.size(); b++) {
        if (b > 64) bad("too many header continuation blocks");
        const auto blk = blocks[b];
        f.need(blk.first, blk.second);
        parse_messages(f, blk.first, blk.first + blk.second, ds, blocks);
    }
    return ds;
}

// Walks a version-1 group B-tree; collects (name, object header address).
inline void walk_group(const File &f, uint64_t node, uint64_t heap_data, int depth,
                std::vector<std::pair<std::string, uint64_t>> &out) {
    if (depth > 32) bad("group B-tree too deep");
    if (!f.sig(node, "TREE") || f.u(node + 4, 1) != 0) bad("expected a group B-tree node");
    const int level = (int)f.u(node + 5, 1);
    const int used = (int)f.u(node + 6, 2);
    uint64_t p = node + 8 + 2 * (uint64_t)f.so; // after siblings
    p += f.sl;                                  // key 0
    for (int k = 0; k < used; k++) {
        const uint64_t child = f.off(p);
        p += f.so + f.sl;
        if (level > 0) {
            walk_group(f, child, heap_data, depth + 1, out);
            continue;
        }
        if (!f.sig(child, "SNOD")) bad("expected a symbol table node");
        const int nsym = (int)f.u(child + 6, 2);
        const uint64_t esz = 2 * (uint64_t)f.so + 24;
        for (int e = 0; e < nsym; e++) {
            const uint64_t ent = child + 8 + e * esz;
            const uint64_t name_off = f.off(ent);
            const uint64_t obj = f.off(ent + f.so);
            out.push_back({f.cstr(heap_data + name_off, 1024), obj});
        }
    }
}

inline void unshuffle(std::vector<uint8_t> &buf, size_t elem) {
    if (elem <= 1) return;
    const size_t n = buf.size() / elem;
    std::vector<uint8_t> out(buf.size());
    for (size_t b = 0; b < elem; b++)
        for (size_t i = 0; i < n; i++) out[i * elem + b] = buf[b * n + i];
    // trailing bytes that do not fill an element stay in place
    for (size_t k = n * elem; k < buf.size(); k++) out[k] = buf[k];
    buf.swap(out);
}

inline std::vector<uint8_t> decode_chunk(const File &f, uint64_t addr, uint64_t nbytes, uint32_t mask, const Dataset &ds,
                                  uint64_t raw_size) {
    f.need(addr, nbytes);
    std::vector<uint8_t> buf(f.d.begin() + (long)addr, f.d.begin() + (long)(addr + nbytes));
    // filters are applied in reverse order on read
    for (int k = (int)ds.filters.size() - 1; k >= 0; k--) {
        if (mask & (1u << k)) continue;
        const Filter &flt = ds.filters[k];
        if (flt.id == 1) { // deflate
            std::vector<uint8_t> out(raw_size);
            uLongf dlen = (uLongf)raw_size;
            const int rc = uncompress(out.data(), &dlen, buf.data(), (uLong)buf.size());
            if (rc != Z_OK) bad("deflate stream could not be decoded");
            out.resize(dlen);
            buf.swap(out);
        } else if (flt.id == 2) { // shuffle
            unshuffle(buf, flt.cd.empty() ? (size_t)ds.type.size : flt.cd[0]);
        } else if (flt.id == 3) { // fletcher32: checksum in the last 4 bytes
            if (buf.size() < 4) bad("short fletcher32 chunk");
            buf.resize(buf.size() - 4);
        } else {
            bad("unsupported filter id " + std::to_string(flt.id));
        }
    }
    return buf;
}

inline void collect_chunks(const File &f, uint64_t node, const Dataset &ds, std::vector<uint8_t> &out, uint64_t n_elem,
                    int depth) {
    if (depth > 32) bad("chunk B-tree too deep");
    if (!f.sig(node, "TREE") || f.u(node + 4, 1) != 1) bad("expected a chunk B-tree node");
    const int level = (int)f.u(node + 5, 1);
    const int used = (int)f.u(node + 6, 2);
    const int nd = (int)ds.chunk.size();
    const uint64_t ksz = 8 + 8 * (uint64_t)nd;
    uint64_t p = node + 8 + 2 * (uint64_t)f.so;
    const uint64_t esz = (uint64_t)ds.type.size;
    const uint64_t chunk_elems = ds.chunk[0];
    for (int k = 0; k < used; k++) {
        const uint64_t csize = f.u(p, 4);
        const uint32_t mask = (uint32_t)f.u(p + 4, 4);
        const uint64_t start = f.u(p + 8, 8); // element offset along dimension 0
        const uint64_t child = f.off(p + ksz);
        p += ksz + f.so;
        if (level > 0) {
            collect_chunks(f, child, ds, out, n_elem, depth + 1);
            continue;
        }
        if (start >= n_elem) bad("chunk outside the dataset");
        const std::vector<uint8_t> raw = decode_chunk(f, child, csize, mask, ds, chunk_elems * esz);
        const uint64_t take = std::min<uint64_t>(chunk_elems, n_elem - start) * esz;
        if (raw.size() < take) bad("decoded chunk shorter than its extent");
        std::memcpy(out.data() + start * esz, raw.data(), take);
    }
}

// Reads a 1-D dataset of 8-byte little-endian elements into `out` bytes.
inline std::vector<uint8_t> read_1d(const File &f, const Dataset &ds) {
    if (ds.dims.size() != 1) bad("dataset is not one-dimensional");
    if (ds.type.size != 8 || ds.type.big_endian || (ds.type.cls != 0 && ds.type.cls != 1))
        bad("dataset elements are not 8-byte little-endian numbers");
    const uint64_t n = ds.dims[0];
    // 2^37 elements (1 TiB) bounds the byte count well inside 64 bits
    if (n > (1ull << 37)) bad("dataset too large");
    if (ds.layout_class == 1 && ds.addr != kUndef) f.need(ds.addr, n * 8);
    std::vector<uint8_t> out(n * 8, 0);
    if (ds.layout_class == 1) {
        if (ds.addr == kUndef) return out; // never written: fill value 0
        std::memcpy(out.data(), &f.d[ds.addr], n * 8);
    } else if (ds.layout_class == 2) {
        if (ds.chunk.size() != 2 || ds.chunk[1] != 8) bad("unexpected chunk shape");
        if (ds.addr != kUndef) collect_chunks(f, ds.addr, ds, out, n, 0);
    } else {
        bad("dataset has no data layout message");
    }
    return out;
}

inline File load(const char *path) {
    File f;
    FILE *fp = std::fopen(path, "rb");
    if (!fp) bad(std::string("cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long sz = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (sz <= 0) {
        std::fclose(fp);
        bad("empty file");
    }
    f.d.resize((size_t)sz);
    const size_t got = std::fread(f.d.data(), 1, (size_t)sz, fp);
    std::fclose(fp);
    if (got != (size_t)sz) bad("short read");
    static const uint8_t kSig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (f.d.size() < 96 || std::memcmp(f.d.data(), kSig, 8) != 0) bad("not an HDF5 file (no signature at offset 0)");
    return f;
}

// Opens a file and lists the members of its root group.
struct Root {
    File f;
    std::vector<std::pair<std::string, uint64_t>> members;

    Dataset dataset(const std::string &name) const {
        for (auto &m : members)
            if (m.first == name) return parse_object(f, m.second);
        bad("dataset \"" + name + "\" not found");
    }
};

inline Root open_root(const char *path) {
    Root r;
    r.f = load(path);
    File &f = r.f;
    const int sb_ver = (int)f.u(8, 1);
    if (sb_ver != 0 && sb_ver != 1) bad("only version 0/1 superblocks are supported");
    f.so = (int)f.u(13, 1);
    f.sl = (int)f.u(14, 1);
    if (f.so != 8 || f.sl != 8) bad("only 8-byte offsets and lengths are supported");
    const uint64_t root_entry = 24 + (sb_ver == 1 ? 4 : 0) + 4 * 8;
    const uint64_t root_obj = f.off(root_entry + 8);
    uint64_t btree = kUndef, heap = kUndef;
    if (f.u(root_entry + 16, 4) == 1) {
        btree = f.off(root_entry + 24);
        heap = f.off(root_entry + 32);
    } else {
        // symbol table message (0x0011) of the root object header
        if (f.u(root_obj, 1) != 1) bad("only version 1 object headers are supported");
        const uint64_t hsize = f.u(root_obj + 8, 4);
        uint64_t p = root_obj + 16, end = p + hsize;
        while (p + 8 <= end) {
            const int type = (int)f.u(p, 2);
            const uint64_t size = f.u(p + 2, 2);
            if (type == 0x0011) {
                btree = f.off(p + 8);
                heap = f.off(p + 8 + f.so);
            }
            p += 8 + size;
        }
        if (btree == kUndef) bad("root group has no symbol table");
    }
    if (!f.sig(heap, "HEAP")) bad("expected a local heap");
    const uint64_t heap_data = f.off(heap + 8 + 2 * (uint64_t)f.sl);
    walk_group(f, btree, heap_data, 0, r.members);
    return r;
}

// A scalar attribute of a dataset as double (IEEE 8 bytes, or a 4/8-byte
// integer of either sign); false if absent.
inline bool attr_number(const Dataset &ds, const char *name, double &out) {
    for (const auto &a : ds.attrs) {
        if (a.name != name) continue;
        if (a.type.big_endian || (int)a.data.size() < a.type.size) bad(std::string("attribute ") + name + " unreadable");
        if (a.type.cls == 1 && a.type.size == 8) {
            std::memcpy(&out, a.data.data(), 8);
            return true;
        }
        if (a.type.cls == 0 && (a.type.size == 4 || a.type.size == 8)) {
            const bool is_signed = a.type.is_signed;
            if (a.type.size == 4) {
                uint32_t v;
                std::memcpy(&v, a.data.data(), 4);
                out = is_signed ? (double)(int32_t)v : (double)v;
            } else {
                uint64_t v;
                std::memcpy(&v, a.data.data(), 8);
                out = is_signed ? (double)(int64_t)v : (double)v;
            }
            return true;
        }
        bad(std::string("attribute ") + name + " has an unsupported type");
    }
    return false;
}

} // namespace h5
} // namespace icempm
