// h5_snapshot.cpp — SnapshotManager::SaveSnapshot / ReadSnapshot
// (snapshotmanager.cpp:299-405) without an HDF5 library: the "d%05d.h5" file
// holds a dataset "Indenter" (IndenterSubdivisions*GridZ*3 doubles) and a
// dataset "Points" (the whole HostSideSOA buffer, capacity*28 doubles, line
// after line) carrying the attributes HSSOA_capacity and HSSOA_size
// (unsigned int) and the SimParams3D values of
// SimParams3D::SaveParametersAsAttributes (parameters_sim_3d.cpp:178-282).
//
// The writer emits a version 0 superblock, a root group with a symbol table
// (one B-tree leaf, a local heap, one symbol node) and two datasets with
// version 1 object headers, in the structures of the published HDF5 file
// format specification and with libhdf5's default node sizes (group leaf K 4,
// group internal K 16, chunk B-tree K 32). "Points" is contiguous, or, with
// compress, chunked in min(2^20, size) elements and deflated at level 5 as
// SaveSnapshot does. Reading goes through h5_file.h.
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "h5_file.h"
#include "icempm_3dm.h"

namespace icempm {
int set_error(int code, const std::string &msg); // capi.hip
}

using icempm::set_error;
using namespace icempm::h5;

namespace {

constexpr uint64_t kUndefAddr = ~0ull;
constexpr int kGroupLeafK = 4, kGroupInternalK = 16, kChunkK = 32;

struct Buf {
    std::vector<uint8_t> b;
    size_t size() const { return b.size(); }
    void u(uint64_t v, int n) {
        for (int k = 0; k < n; k++) b.push_back((uint8_t)(v >> (8 * k)));
    }
    void bytes(const void *p, size_t n) {
        const uint8_t *q = static_cast<const uint8_t *>(p);
        b.insert(b.end(), q, q + n);
    }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad8() { zeros((8 - b.size() % 8) % 8); }
    void put(size_t at, uint64_t v, int n) {
        for (int k = 0; k < n; k++) b[at + k] = (uint8_t)(v >> (8 * k));
    }
};

// Datatype message bodies (class and version 1 in the first byte).
void dt_double(Buf &m) {
    m.u(0x11, 1);
    m.u(0x20, 1); // little-endian, implied mantissa msb
    m.u(0x3f, 1); // sign bit at 63
    m.u(0x00, 1);
    m.u(8, 4);
    m.u(0, 2);    // bit offset
    m.u(64, 2);   // precision
    m.u(52, 1);   // exponent location
    m.u(11, 1);   // exponent size
    m.u(0, 1);    // mantissa location
    m.u(52, 1);   // mantissa size
    m.u(1023, 4); // exponent bias
}
void dt_int32(Buf &m, bool is_signed) {
    m.u(0x10, 1);
    m.u(is_signed ? 0x08 : 0x00, 1);
    m.u(0, 2);
    m.u(4, 4);
    m.u(0, 2);
    m.u(32, 2);
}
// Simple dataspace, version 1, rank 1, with maximum dimensions.
void ds_rank1(Buf &m, uint64_t n) {
    m.u(1, 1);
    m.u(1, 1);
    m.u(1, 1);
    m.u(0, 1);
    m.u(0, 4);
    m.u(n, 8);
    m.u(n, 8);
}

struct Message {
    int type;
    Buf body;
};

void add_attr(std::vector<Message> &msgs, const char *name, int kind, const void *value) {
    // kind 0 double, 1 int32, 2 uint32
    Buf dt, sp;
    if (kind == 0) dt_double(dt);
    else dt_int32(dt, kind == 1);
    ds_rank1(sp, 1);
    Message m{0x000C, {}};
    const size_t nlen = std::strlen(name) + 1;
    m.body.u(1, 1);
    m.body.u(0, 1);
    m.body.u(nlen, 2);
    m.body.u(dt.size(), 2);
    m.body.u(sp.size(), 2);
    m.body.bytes(name, nlen);
    m.body.pad8();
    m.body.bytes(dt.b.data(), dt.size());
    m.body.pad8();
    m.body.bytes(sp.b.data(), sp.size());
    m.body.pad8();
    m.body.bytes(value, kind == 0 ? 8 : 4);
    m.body.pad8();
    msgs.push_back(std::move(m));
}

// Version 1 object header: 16-byte prefix, then the messages (8-aligned).
std::vector<uint8_t> object_header(const std::vector<Message> &msgs) {
    Buf h;
    size_t total = 0;
    for (const auto &m : msgs) total += 8 + (m.body.size() + 7) / 8 * 8;
    h.u(1, 1);
    h.u(0, 1);
    h.u(msgs.size(), 2);
    h.u(1, 4);
    h.u(total, 4);
    h.u(0, 4);
    for (const auto &m : msgs) {
        const size_t sz = (m.body.size() + 7) / 8 * 8;
        h.u(m.type, 2);
        h.u(sz, 2);
        h.u(m.type == 0x0003 || m.type == 0x0005 || m.type == 0x000B ? 1 : 0, 1); // constant messages
        h.zeros(3);
        h.bytes(m.body.b.data(), m.body.size());
        h.zeros(sz - m.body.size());
    }
    return h.b;
}

Message msg_dataspace(uint64_t n) {
    Message m{0x0001, {}};
    ds_rank1(m.body, n);
    return m;
}
Message msg_double_type() {
    Message m{0x0003, {}};
    dt_double(m.body);
    return m;
}
Message msg_fill(int alloc_time) {
    Message m{0x0005, {}};
    m.body.u(2, 1);          // version 2
    m.body.u(alloc_time, 1); // 2 late (contiguous), 3 incremental (chunked)
    m.body.u(2, 1);          // write fill value if set by the user
    m.body.u(1, 1);          // fill value defined (the default, zero)
    m.body.u(0, 4);
    return m;
}

struct ParamAttr {
    const char *name;
    int kind;
    size_t offset;
};

#define PD(f) {#f, 0, offsetof(icempm_SimParams3D, f)}
#define PI(f) {#f, 1, offsetof(icempm_SimParams3D, f)}
// SaveParametersAsAttributes (parameters_sim_3d.cpp:183-281), in its order.
const ParamAttr kParams[] = {
    PI(SetupType), PI(nPtsTotal), PI(GridXTotal), PI(GridY), PI(GridZ), PD(cellsize), PI(IndenterSubdivisions),
    PD(InitialTimeStep), PD(SimulationEndTime), PD(SimulationTime), PI(UpdateEveryNthStep), PI(SnapshotPeriod),
    PI(SimulationStep), PD(Gravity), PD(Density), PD(PoissonsRatio), PD(YoungsModulus), PD(IceCompressiveStrength),
    PD(IceTensileStrength), PD(IceTensileStrength2), PD(IceShearStrength), PD(DP_tan_phi), PD(DP_threshold_p),
    PD(GrainVariability), PD(indenter_x), PD(indenter_y), PD(indenter_x_initial), PD(indenter_y_initial),
    PD(IndDiameter), PD(IndVelocity), PD(ParticleVolume), PD(ParticleMass), PD(Volume), PI(nPartitions),
    PI(GridHaloSize), PI(PointTransferFrequency), PD(ExtraSpaceForIncomingPoints), PD(PointsTransferBufferFraction),
    PD(RebalanceThresholdFreeSpaceRemaining), PD(RebalanceThresholdDisabledPercentage), PD(animation_threshold_pos),
};
#undef PD
#undef PI

struct Chunk {
    std::vector<uint8_t> data;
    uint64_t offset; // first element
};

// Version 1 B-tree of chunks (node type 1) for a 1-D dataset of doubles,
// appended at `base` of the file image; returns the root address.
uint64_t write_chunk_btree(Buf &file, const std::vector<uint64_t> &chunk_addr, const std::vector<Chunk> &chunks,
                           uint64_t chunk_elems) {
    const int fan = 2 * kChunkK;
    const uint64_t key_size = 4 + 4 + 2 * 8;
    const uint64_t node_size = 24 + (uint64_t)fan * 8 + ((uint64_t)fan + 1) * key_size;
    struct Entry {
        uint64_t addr, size, offset;
    };
    std::vector<Entry> level_entries;
    for (size_t k = 0; k < chunks.size(); k++) level_entries.push_back({chunk_addr[k], chunks[k].data.size(), chunks[k].offset});
    const uint64_t end_offset = chunks.back().offset + chunk_elems;
    int level = 0;
    while (true) {
        std::vector<Entry> parents;
        for (size_t s = 0; s < level_entries.size(); s += fan) {
            const size_t used = std::min<size_t>(fan, level_entries.size() - s);
            const uint64_t at = file.size();
            file.bytes("TREE", 4);
            file.u(1, 1);
            file.u(level, 1);
            file.u(used, 2);
            file.u(kUndefAddr, 8);
            file.u(kUndefAddr, 8);
            for (size_t k = 0; k < used; k++) {
                const Entry &e = level_entries[s + k];
                file.u(e.size, 4);
                file.u(0, 4);
                file.u(e.offset, 8);
                file.u(0, 8);
                file.u(e.addr, 8);
            }
            // right key: just past the last chunk of this node
            const uint64_t right = (s + used < level_entries.size()) ? level_entries[s + used].offset : end_offset;
            file.u(0, 4);
            file.u(0, 4);
            file.u(right, 8);
            file.u(0, 8);
            file.zeros(node_size - (file.size() - at));
            parents.push_back({at, level_entries[s].size, level_entries[s].offset});
        }
        if (parents.size() == 1) return parents[0].addr;
        level_entries.swap(parents);
        level++;
    }
}

} // namespace

extern "C" int icempm_save_snapshot(const char *path, const icempm_SimParams3D *prms, const double *hssoa,
                                    unsigned capacity, unsigned size, const double *indenter_bins, int compress) {
    if (!path || !prms || !hssoa || size > capacity)
        return set_error(ICEMPM_ERR_INVALID_ARG, "save_snapshot: bad arguments");
    const long long nind = 3LL * prms->IndenterSubdivisions * prms->GridZ;
    if (nind < 0 || (nind > 0 && !indenter_bins))
        return set_error(ICEMPM_ERR_INVALID_ARG, "save_snapshot: indenter accumulators missing");
    try {
        const uint64_t npts = (uint64_t)capacity * ICEMPM_N_PTS_ARRAYS;
        // ---- the two object headers (their addresses are patched in later)
        std::vector<Message> ind_msgs{msg_dataspace((uint64_t)nind), msg_double_type(), msg_fill(2)};
        Message ind_layout{0x0008, {}};
        ind_layout.body.u(3, 1);
        ind_layout.body.u(1, 1);
        const size_t ind_addr_at = ind_layout.body.size();
        ind_layout.body.u(0, 8);
        ind_layout.body.u((uint64_t)nind * 8, 8);
        ind_msgs.push_back(ind_layout);

        std::vector<Chunk> chunks;
        uint64_t chunk_elems = 0;
        std::vector<Message> pts_msgs{msg_dataspace(npts), msg_double_type(), msg_fill(compress ? 3 : 2)};
        Message pts_layout{0x0008, {}};
        pts_layout.body.u(3, 1);
        size_t pts_addr_at;
        if (compress) {
            if (size == 0) return set_error(ICEMPM_ERR_INVALID_ARG, "save_snapshot: compress needs size > 0");
            chunk_elems = std::min<uint64_t>(1024 * 1024, size);
            for (uint64_t o = 0; o < npts; o += chunk_elems) {
                std::vector<uint8_t> raw(chunk_elems * 8, 0);
                const uint64_t take = std::min<uint64_t>(chunk_elems, npts - o);
                std::memcpy(raw.data(), hssoa + o, take * 8);
                uLongf clen = compressBound((uLong)raw.size());
                Chunk c{std::vector<uint8_t>(clen), o};
                if (compress2(c.data.data(), &clen, raw.data(), (uLong)raw.size(), 5) != Z_OK)
                    return set_error(ICEMPM_ERR_INVALID_ARG, "save_snapshot: deflate failed");
                c.data.resize(clen);
                chunks.push_back(std::move(c));
            }
            pts_layout.body.u(2, 1);
            pts_layout.body.u(2, 1);
            pts_addr_at = pts_layout.body.size();
            pts_layout.body.u(0, 8);
            pts_layout.body.u(chunk_elems, 4);
            pts_layout.body.u(8, 4);
        } else {
            pts_layout.body.u(1, 1);
            pts_addr_at = pts_layout.body.size();
            pts_layout.body.u(0, 8);
            pts_layout.body.u(npts * 8, 8);
        }
        pts_msgs.push_back(pts_layout);
        if (compress) {
            Message flt{0x000B, {}};
            flt.body.u(1, 1);
            flt.body.u(1, 1);
            flt.body.zeros(6);
            flt.body.u(1, 2);  // deflate
            flt.body.u(8, 2);  // name length
            flt.body.u(1, 2);  // optional
            flt.body.u(1, 2);  // one client value
            flt.body.bytes("deflate\0", 8);
            flt.body.u(5, 4);  // level
            flt.body.u(0, 4);  // padding to an even count
            pts_msgs.push_back(flt);
        }
        const unsigned cap_u = capacity, size_u = size;
        add_attr(pts_msgs, "HSSOA_capacity", 2, &cap_u);
        add_attr(pts_msgs, "HSSOA_size", 2, &size_u);
        for (const ParamAttr &pa : kParams)
            add_attr(pts_msgs, pa.name, pa.kind, reinterpret_cast<const uint8_t *>(prms) + pa.offset);

        // ---- file image: superblock, root group, heap, symbol node, headers
        Buf f;
        f.bytes("\x89HDF\r\n\x1a\n", 8);
        f.u(0, 1);
        f.u(0, 1);
        f.u(0, 1);
        f.u(0, 1);
        f.u(0, 1);
        f.u(8, 1);
        f.u(8, 1);
        f.u(0, 1);
        f.u(kGroupLeafK, 2);
        f.u(kGroupInternalK, 2);
        f.u(0, 4);
        f.u(0, 8);          // base address
        f.u(kUndefAddr, 8); // free-space info
        const size_t eof_at = f.size();
        f.u(0, 8);          // end of file (patched)
        f.u(kUndefAddr, 8); // driver info
        // root symbol table entry: name offset, header address, cache type 1
        const uint64_t root_hdr = 96, btree = 136;
        const uint64_t btree_size = 24 + (uint64_t)(2 * kGroupInternalK) * 8 + (uint64_t)(2 * kGroupInternalK + 1) * 8;
        const uint64_t heap = btree + btree_size;
        const uint64_t heap_data = heap + 32;
        // heap data: "" | "Indenter" | "Points", each null-terminated, 8-aligned
        Buf names;
        names.zeros(8);
        const uint64_t off_ind = names.size();
        names.bytes("Indenter", 9);
        names.pad8();
        const uint64_t off_pts = names.size();
        names.bytes("Points", 7);
        names.pad8();
        const uint64_t snod = heap_data + names.size();
        const uint64_t snod_size = 8 + (uint64_t)(2 * kGroupLeafK) * (2 * 8 + 24);
        f.u(0, 8);
        f.u(root_hdr, 8);
        f.u(1, 4);
        f.u(0, 4);
        f.u(btree, 8);
        f.u(heap, 8);
        // root object header: one symbol table message
        f.u(1, 1);
        f.u(0, 1);
        f.u(1, 2);
        f.u(1, 4);
        f.u(24, 4);
        f.u(0, 4);
        f.u(0x0011, 2);
        f.u(16, 2);
        f.u(0, 1);
        f.zeros(3);
        f.u(btree, 8);
        f.u(heap, 8);
        // group B-tree leaf: key 0 (""), the symbol node, key 1 (last name)
        f.bytes("TREE", 4);
        f.u(0, 1);
        f.u(0, 1);
        f.u(1, 2);
        f.u(kUndefAddr, 8);
        f.u(kUndefAddr, 8);
        f.u(0, 8);
        f.u(snod, 8);
        f.u(off_pts, 8);
        f.zeros(heap - f.size());
        // local heap (no free block: offset 1 is the null free-list mark)
        f.bytes("HEAP", 4);
        f.u(0, 1);
        f.zeros(3);
        f.u(names.size(), 8);
        f.u(1, 8);
        f.u(heap_data, 8);
        f.bytes(names.b.data(), names.size());
        // symbol node with the two datasets (patched addresses)
        f.bytes("SNOD", 4);
        f.u(1, 1);
        f.u(0, 1);
        f.u(2, 2);
        const size_t ent0 = f.size();
        f.zeros(snod_size - 8);
        // object headers
        std::vector<uint8_t> ind_hdr = object_header(ind_msgs);
        std::vector<uint8_t> pts_hdr = object_header(pts_msgs);
        const uint64_t ind_hdr_addr = f.size();
        f.bytes(ind_hdr.data(), ind_hdr.size());
        const uint64_t pts_hdr_addr = f.size();
        f.bytes(pts_hdr.data(), pts_hdr.size());
        f.put(ent0, off_ind, 8);
        f.put(ent0 + 8, ind_hdr_addr, 8);
        f.put(ent0 + 40, off_pts, 8);
        f.put(ent0 + 48, pts_hdr_addr, 8);
        // raw data
        f.pad8();
        const uint64_t ind_data = f.size();
        if (nind > 0) f.bytes(indenter_bins, (size_t)nind * 8);
        // layout body offset of the address inside the header image: prefix 16,
        // then each message 8 + padded body
        auto patch_layout = [&](uint64_t hdr_addr, const std::vector<Message> &msgs, size_t body_off, uint64_t v) {
            uint64_t p = hdr_addr + 16;
            for (const auto &m : msgs) {
                if (m.type == 0x0008) {
                    f.put(p + 8 + body_off, v, 8);
                    return;
                }
                p += 8 + (m.body.size() + 7) / 8 * 8;
            }
        };
        patch_layout(ind_hdr_addr, ind_msgs, ind_addr_at, nind > 0 ? ind_data : kUndefAddr);
        f.pad8();
        if (compress) {
            std::vector<uint64_t> addr;
            for (const auto &c : chunks) {
                addr.push_back(f.size());
                f.bytes(c.data.data(), c.data.size());
            }
            f.pad8();
            const uint64_t root = write_chunk_btree(f, addr, chunks, chunk_elems);
            patch_layout(pts_hdr_addr, pts_msgs, pts_addr_at, root);
        } else {
            const uint64_t pts_data = f.size();
            f.bytes(hssoa, (size_t)npts * 8);
            patch_layout(pts_hdr_addr, pts_msgs, pts_addr_at, pts_data);
        }
        f.put(eof_at, f.size(), 8);

        const std::string tmp = std::string(path) + ".tmp";
        FILE *fp = std::fopen(tmp.c_str(), "wb");
        if (!fp) return set_error(ICEMPM_ERR_INVALID_ARG, std::string("save_snapshot: cannot create ") + tmp);
        const size_t wrote = std::fwrite(f.b.data(), 1, f.size(), fp);
        const int closed = std::fclose(fp);
        if (wrote != f.size() || closed != 0 || std::rename(tmp.c_str(), path) != 0) {
            std::remove(tmp.c_str());
            return set_error(ICEMPM_ERR_INVALID_ARG, std::string("save_snapshot: cannot write ") + path);
        }
        return ICEMPM_OK;
    } catch (const std::bad_alloc &) {
        return set_error(ICEMPM_ERR_CAPACITY, "save_snapshot: out of host memory");
    } catch (const std::exception &e) {
        // no exception may cross the C boundary (std::length_error of a size
        // taken from a damaged file, for one)
        return set_error(ICEMPM_ERR_INVALID_ARG, std::string("save_snapshot: ") + e.what());
    }
}

extern "C" int icempm_read_snapshot(const char *path, icempm_SimParams3D *prms, double *hssoa, unsigned capacity,
                                    unsigned *capacity_out, unsigned *size_out, double *indenter_bins,
                                    int n_indenter_bins) {
    if (!path) return set_error(ICEMPM_ERR_INVALID_ARG, "read_snapshot: null path");
    try {
        const Root root = open_root(path);
        const Dataset pts = root.dataset("Points");
        // counts are whole numbers in [0, 2^32); checked before conversion
        const auto count_attr = [&](const char *name) -> uint64_t {
            double c = 0;
            if (!attr_number(pts, name, c)) bad(std::string("Points has no ") + name);
            if (!(c >= 0 && c <= 4294967295.0)) bad(std::string("Points attribute ") + name + " out of range");
            return (uint64_t)c;
        };
        const uint64_t file_cap = count_attr("HSSOA_capacity");
        const uint64_t file_size = count_attr("HSSOA_size");
        if (file_size > file_cap || file_cap > 0xffffffffull) bad("inconsistent HSSOA_capacity / HSSOA_size");
        if (pts.dims.size() != 1 || pts.dims[0] != file_cap * ICEMPM_N_PTS_ARRAYS)
            bad("Points does not hold capacity x 28 values");
        if (capacity_out) *capacity_out = (unsigned)file_cap;
        if (size_out) *size_out = (unsigned)file_size;
        if (prms) {
            // SimParams3D::ReadParametersFromAttributes (parameters_sim_3d.cpp:285-372):
            // Reset, the saved values, DomainDimensionX, then the derived constants.
            // DP_threshold_p is read back too (the reference opens but never
            // reads it and falls back to the default).
            icempm_params_reset(prms);
            for (const ParamAttr &pa : kParams) {
                double x;
                const bool optional = std::strcmp(pa.name, "animation_threshold_pos") == 0;
                if (!attr_number(pts, pa.name, x)) {
                    if (optional) continue;
                    bad(std::string("Points has no attribute ") + pa.name);
                }
                uint8_t *dst = reinterpret_cast<uint8_t *>(prms) + pa.offset;
                if (pa.kind == 0) std::memcpy(dst, &x, 8);
                else {
                    if (!(x >= -2147483648.0 && x <= 2147483647.0))
                        bad(std::string("Points attribute ") + pa.name + " is not a 32-bit integer");
                    const int32_t iv = (int32_t)x;
                    std::memcpy(dst, &iv, 4);
                }
            }
            prms->DomainDimensionX = prms->GridXTotal * prms->cellsize;
            icempm_params_compute_camclay(prms);
            icempm_params_compute_helper_variables(prms);
        }
        if (!hssoa) return ICEMPM_OK; // size query
        if (capacity < file_cap) return set_error(ICEMPM_ERR_CAPACITY, "read_snapshot: buffer smaller than HSSOA_capacity");
        const std::vector<uint8_t> bytes = read_1d(root.f, pts);
        const double *src = reinterpret_cast<const double *>(bytes.data());
        for (int l = 0; l < ICEMPM_N_PTS_ARRAYS; l++) {
            std::memcpy(hssoa + (size_t)l * capacity, src + (size_t)l * file_cap, file_cap * 8);
            if (capacity > file_cap) std::memset(hssoa + (size_t)l * capacity + file_cap, 0, (capacity - file_cap) * 8);
        }
        if (indenter_bins) {
            const Dataset ind = root.dataset("Indenter");
            if (ind.dims.size() != 1) bad("Indenter is not one-dimensional");
            if (ind.dims[0] > (uint64_t)std::max(n_indenter_bins, 0))
                return set_error(ICEMPM_ERR_CAPACITY, "read_snapshot: Indenter holds more values than the buffer");
            if (ind.dims[0] > 0) {
                const std::vector<uint8_t> ib = read_1d(root.f, ind);
                std::memcpy(indenter_bins, ib.data(), ib.size());
            }
        }
        return ICEMPM_OK;
    } catch (const H5Error &e) {
        return set_error(ICEMPM_ERR_INVALID_ARG, std::string("read_snapshot: ") + path + ": " + e.msg);
    } catch (const std::bad_alloc &) {
        return set_error(ICEMPM_ERR_CAPACITY, "read_snapshot: out of host memory");
    } catch (const std::exception &e) {
        // no exception may cross the C boundary (std::length_error of a size
        // taken from a damaged file, for one)
        return set_error(ICEMPM_ERR_INVALID_ARG, std::string("read_snapshot: ") + e.what());
    }
}
