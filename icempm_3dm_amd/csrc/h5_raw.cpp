// h5_raw.cpp — reader of the raw point-cloud files that
// SnapshotManager::LoadRawPoints opens (snapshotmanager.cpp:20-47): an HDF5
// file with 1-D datasets "x", "y", "z" (doubles), "llGrainIDs" (64-bit
// integers) and a double attribute "volume" on "llGrainIDs".
//
// The image has no HDF5 library, so this is a small reader of the subset of
// the HDF5 file format (version 0/1 superblock, version 1 object headers,
// symbol-table groups, contiguous or chunked layout version 3, deflate and
// shuffle filters, little-endian fixed-point and IEEE types) that such files
// use, written from the published HDF5 file format specification. Every read
// is bounds-checked against the file size; anything outside the subset is an
// error, never a guess.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "icempm_3dm.h"

namespace icempm {
int set_error(int code, const std::string &msg); // capi.hip
}

namespace {

using icempm::set_error;

struct H5Error {
    std::string msg;
};

[[noreturn]] void bad(const std::string &m) { throw H5Error{m}; }

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

Datatype parse_datatype(const File &f, uint64_t at) {
    Datatype t;
    const uint64_t b0 = f.u(at, 1);
    t.cls = (int)(b0 & 0x0f);
    const uint64_t bits = f.u(at + 1, 3);
    t.size = (int)f.u(at + 4, 4);
    if (t.cls == 0 || t.cls == 1) {
        t.big_endian = (bits & 1) != 0;
        if (t.cls == 1 && (bits & 0x40)) bad("VAX floating-point order is not supported");
    }
    return t;
}

std::vector<uint64_t> parse_dataspace(const File &f, uint64_t at) {
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

void parse_filters(const File &f, uint64_t at, Dataset &ds) {
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

void parse_attribute(const File &f, uint64_t at, Dataset &ds) {
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
    for (uint64_t d : dims) count *= d;
    const uint64_t nbytes = count * (uint64_t)a.type.size;
    f.need(p, nbytes);
    a.data.assign(f.d.begin() + (long)p, f.d.begin() + (long)(p + nbytes));
    ds.attrs.push_back(a);
}

void parse_messages(const File &f, uint64_t p, uint64_t end, Dataset &ds, std::vector<std::pair<uint64_t, uint64_t>> &more) {
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

Dataset parse_object(const File &f, uint64_t addr) {
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
void walk_group(const File &f, uint64_t node, uint64_t heap_data, int depth,
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

void unshuffle(std::vector<uint8_t> &buf, size_t elem) {
    if (elem <= 1) return;
    const size_t n = buf.size() / elem;
    std::vector<uint8_t> out(buf.size());
    for (size_t b = 0; b < elem; b++)
        for (size_t i = 0; i < n; i++) out[i * elem + b] = buf[b * n + i];
    // trailing bytes that do not fill an element stay in place
    for (size_t k = n * elem; k < buf.size(); k++) out[k] = buf[k];
    buf.swap(out);
}

std::vector<uint8_t> decode_chunk(const File &f, uint64_t addr, uint64_t nbytes, uint32_t mask, const Dataset &ds,
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

void collect_chunks(const File &f, uint64_t node, const Dataset &ds, std::vector<uint8_t> &out, uint64_t n_elem,
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
std::vector<uint8_t> read_1d(const File &f, const Dataset &ds) {
    if (ds.dims.size() != 1) bad("dataset is not one-dimensional");
    if (ds.type.size != 8 || ds.type.big_endian || (ds.type.cls != 0 && ds.type.cls != 1))
        bad("dataset elements are not 8-byte little-endian numbers");
    const uint64_t n = ds.dims[0];
    std::vector<uint8_t> out(n * 8, 0);
    if (ds.layout_class == 1) {
        if (ds.addr == kUndef) return out; // never written: fill value 0
        f.need(ds.addr, n * 8);
        std::memcpy(out.data(), &f.d[ds.addr], n * 8);
    } else if (ds.layout_class == 2) {
        if (ds.chunk.size() != 2 || ds.chunk[1] != 8) bad("unexpected chunk shape");
        if (ds.addr != kUndef) collect_chunks(f, ds.addr, ds, out, n, 0);
    } else {
        bad("dataset has no data layout message");
    }
    return out;
}

File load(const char *path) {
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

} // namespace

extern "C" int icempm_read_raw_points(const char *path, double *hssoa, unsigned capacity, int *n_points_out,
                                      double *volume_out) {
    if (!path || !n_points_out) return set_error(ICEMPM_ERR_INVALID_ARG, "read_raw_points: null argument");
    try {
        File f = load(path);
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
        std::vector<std::pair<std::string, uint64_t>> members;
        walk_group(f, btree, heap_data, 0, members);
        auto find = [&](const char *name) -> Dataset {
            for (auto &m : members)
                if (m.first == name) return parse_object(f, m.second);
            bad(std::string("dataset \"") + name + "\" not found");
        };
        const Dataset grains = find("llGrainIDs");
        if (grains.dims.size() != 1) bad("llGrainIDs is not one-dimensional");
        const uint64_t n = grains.dims[0];
        if (n > 0x7fffffffull) bad("too many points");
        *n_points_out = (int)n;
        double volume = 0.0;
        bool have_volume = false;
        for (const auto &a : grains.attrs)
            if (a.name == "volume" && a.type.cls == 1 && a.type.size == 8 && a.data.size() >= 8) {
                std::memcpy(&volume, a.data.data(), 8);
                have_volume = true;
            }
        if (!have_volume) bad("llGrainIDs has no double attribute \"volume\"");
        if (volume_out) *volume_out = volume;
        if (!hssoa) return ICEMPM_OK; // size query
        if (n > capacity) return set_error(ICEMPM_ERR_CAPACITY, "read_raw_points: file holds more points than capacity");
        const char *axes[3] = {"x", "y", "z"};
        for (int a = 0; a < 3; a++) {
            const Dataset ds = find(axes[a]);
            if (ds.type.cls != 1) bad(std::string("dataset \"") + axes[a] + "\" is not floating-point");
            if (ds.dims.empty() || ds.dims[0] != n) bad("coordinate and grain datasets differ in length");
            const std::vector<uint8_t> bytes = read_1d(f, ds);
            std::memcpy(hssoa + (size_t)(ICEMPM_POSX + a) * capacity, bytes.data(), bytes.size());
        }
        if (grains.type.cls != 0) bad("llGrainIDs is not an integer dataset");
        const std::vector<uint8_t> gb = read_1d(f, grains);
        // line 0 holds the 64-bit grain ids as stored (snapshotmanager.cpp:40)
        std::memcpy(hssoa + (size_t)ICEMPM_IDX_UTILITY * capacity, gb.data(), gb.size());
        return ICEMPM_OK;
    } catch (const H5Error &e) {
        return set_error(ICEMPM_ERR_INVALID_ARG, std::string("read_raw_points: ") + path + ": " + e.msg);
    } catch (const std::bad_alloc &) {
        return set_error(ICEMPM_ERR_CAPACITY, "read_raw_points: out of host memory");
    }
}

// The geometry half of SnapshotManager::LoadRawPoints (snapshotmanager.cpp:
// 30-107) on points [0, n) of a HostSideSOA buffer that holds the file's
// positions (lines 4-6) and grain ids (line 0), with prms->Volume set:
// block boundaries into xmin..zmax (pre-offset, as the reference stores
// them), SetupType 2 solid/liquid marking against BlockCutout
// (HostSideSOA::MarkSolidAndLiquid, host_side_soa.cpp:33-53), the offset that
// centres the block in x and z and puts its bottom at y = 2h
// (HostSideSOA::offsetBlock, :20-31), a fresh state (the other lines zero,
// then HostSideSOA::InitializeBlock, :88-97: Jp_inv = 1, Fe = I), the
// indenter's starting position, ParticleVolume = Volume / n and
// ComputeHelperVariables. The reference then sorts the points by cell on the
// host (RemoveDisabledAndSort); here the caller does that on the device
// after the upload (icempm_remove_disabled_and_sort).
extern "C" int icempm_prepare_raw_points(icempm_SimParams3D *prms, double *hssoa, unsigned capacity, int n) {
    if (!prms || !hssoa || n <= 0 || (unsigned)n > capacity)
        return set_error(ICEMPM_ERR_INVALID_ARG, "prepare_raw_points: bad arguments");
    double *line[ICEMPM_N_PTS_ARRAYS];
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) line[l] = hssoa + (size_t)l * capacity;
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = hi[k] = line[ICEMPM_POSX + k][0];
        for (int i = 1; i < n; i++) {
            const double v = line[ICEMPM_POSX + k][i];
            if (v < lo[k]) lo[k] = v;
            if (hi[k] < v) hi[k] = v;
        }
    }
    prms->nPtsTotal = n;
    prms->xmin = lo[0];
    prms->ymin = lo[1];
    prms->zmin = lo[2];
    prms->xmax = hi[0];
    prms->ymax = hi[1];
    prms->zmax = hi[2];
    const double h = prms->cellsize;
    const double box_x = prms->GridXTotal * h;
    const double box_z = prms->GridZ * h;
    const double length = prms->xmax - prms->xmin;
    const double width = prms->zmax - prms->zmin;
    const double offset[3] = {(box_x - length) / 2, 2 * h - prms->ymin, (box_z - width) / 2};
    if (prms->SetupType == 2) {
        const double center[3] = {(lo[0] + hi[0]) * 0.5, hi[1] - prms->BlockCutout[1] / 2, (lo[2] + hi[2]) * 0.5};
        double from[3], to[3];
        for (int k = 0; k < 3; k++) {
            from[k] = center[k] - 0.5 * prms->BlockCutout[k];
            to[k] = center[k] + 0.5 * prms->BlockCutout[k];
        }
        for (int i = 0; i < n; i++) {
            bool outside = false;
            for (int k = 0; k < 3 && !outside; k++) outside = line[ICEMPM_POSX + k][i] < from[k];
            for (int k = 0; k < 3 && !outside; k++) outside = line[ICEMPM_POSX + k][i] > to[k];
            if (outside) {
                uint64_t u;
                std::memcpy(&u, &line[ICEMPM_IDX_UTILITY][i], 8);
                u |= (uint64_t)ICEMPM_STATUS_LIQUID;
                std::memcpy(&line[ICEMPM_IDX_UTILITY][i], &u, 8);
            }
        }
    }
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < n; i++) line[ICEMPM_POSX + k][i] += offset[k];
    for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) {
        if (l == ICEMPM_IDX_UTILITY || (l >= ICEMPM_POSX && l < ICEMPM_POSX + 3)) continue;
        std::memset(line[l], 0, sizeof(double) * (size_t)n);
    }
    for (int i = 0; i < n; i++) {
        line[ICEMPM_IDX_JP_INV][i] = 1.0;
        for (int k = 0; k < 3; k++) line[ICEMPM_FE00 + 4 * k][i] = 1.0;
    }
    const double block_left = offset[0];
    const double block_top = prms->ymax + offset[1];
    const double r = prms->IndDiameter / 2;
    const double ht = r - prms->IndDepth;
    const double x_ind_offset = std::sqrt(r * r - ht * ht);
    prms->indenter_x = std::floor((block_left - x_ind_offset) / h) * h;
    if (prms->SetupType == 0) prms->indenter_y = block_top + ht;
    else if (prms->SetupType == 1) prms->indenter_y = std::ceil(block_top / h) * h;
    prms->indenter_x_initial = prms->indenter_x;
    prms->indenter_y_initial = prms->indenter_y;
    prms->ParticleVolume = prms->Volume / n;
    icempm_params_compute_helper_variables(prms);
    return ICEMPM_OK;
}
