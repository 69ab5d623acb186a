// h5_raw.cpp — reader of the raw point-cloud files that
// SnapshotManager::LoadRawPoints opens (snapshotmanager.cpp:20-47): an HDF5
// file with 1-D datasets "x", "y", "z" (doubles), "llGrainIDs" (64-bit
// integers) and a double attribute "volume" on "llGrainIDs".
//
// The image has no HDF5 library; the file is read with the small reader of
// h5_file.h.
#include <cmath>
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

extern "C" int icempm_read_raw_points(const char *path, double *hssoa, unsigned capacity, int *n_points_out,
                                      double *volume_out) {
    if (!path || !n_points_out) return set_error(ICEMPM_ERR_INVALID_ARG, "read_raw_points: null argument");
    try {
        const Root root = open_root(path);
        const File &f = root.f;
        auto find = [&](const char *name) { return root.dataset(name); };
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
    } catch (const std::exception &e) {
        // no exception may cross the C boundary (std::length_error of a size
        // taken from a damaged file, for one)
        return set_error(ICEMPM_ERR_INVALID_ARG, std::string("read_raw_points: ") + e.what());
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
