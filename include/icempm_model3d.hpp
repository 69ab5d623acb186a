/*
 * icempm_model3d.hpp — header-only C++ mirror of the reference's Model3D
 * (simulation/model_3d.{h,cpp}) and of the HostSideSOA buffer it steps
 * (simulation/host_side_soa.{h,cpp}), for one partition, on top of the
 * GPU_Implementation6 mirror.
 *
 * Model3D::Step runs one output cycle exactly as the reference's does
 * (model_3d.cpp:24-112): the substeps up to the next multiple of
 * UpdateEveryNthStep (on the device, without a host synchronization per
 * substep), the copy back with the cycle force and the 1/(N h^2) sensor
 * scaling, one line of the indenter-force log, and the rebalance when the
 * free space or the share of disabled points crosses its threshold. The
 * rebalance's RemoveDisabledAndSort runs on the device; the host buffer is
 * then refreshed from it, where the reference sorts on the host and uploads.
 * The reference's logging, timing log, snapshot mutexes and point transfer
 * between partitions are not part of the step and are left out.
 */
#ifndef ICEMPM_MODEL3D_HPP
#define ICEMPM_MODEL3D_HPP

#include <array>
#include <cmath>
#include <cstring>
#include <vector>

#include "icempm_gpu_implementation6.hpp"

namespace icempm {

// HostSideSOA (host_side_soa.h:51-70): 28 lines of `capacity` doubles.
struct HostSideSOA {
    std::vector<double> host_buffer;
    unsigned capacity = 0, size = 0;

    // host_side_soa.cpp:70-85
    void Allocate(unsigned cap) {
        capacity = cap;
        size = 0;
        host_buffer.assign((size_t)cap * ICEMPM_N_PTS_ARRAYS, 0.0);
    }
    // host_side_soa.h:65
    double *getPointerToLine(int k) { return host_buffer.data() + (size_t)k * capacity; }
    long long utility(unsigned i) const {
        long long u;
        std::memcpy(&u, host_buffer.data() + (size_t)ICEMPM_IDX_UTILITY * capacity + i, sizeof(u));
        return u;
    }
    unsigned count_disabled() const {
        unsigned c = 0;
        for (unsigned i = 0; i < size; i++)
            if (utility(i) & ICEMPM_STATUS_DISABLED) c++;
        return c;
    }
};

class Model3D {
public:
    explicit Model3D(int device = 0) : gpu(device) {}

    icempm_SimParams3D prms{};
    GPU_Implementation6 gpu;
    HostSideSOA hssoa;
    // one entry per cycle, as the reference's "indenter_force" log line
    // (model_3d.cpp:70): time, |F|, F_x, F_y
    std::vector<std::array<double, 4>> indenter_force_log;
    int rebalances = 0;

    // model_3d.cpp:122-130
    void Reset() {
        prms.SimulationStep = 0;
        prms.SimulationTime = 0;
    }
    // SnapshotManager's hand-over of a loaded state to the device
    // (snapshotmanager.cpp:103-106): allocate, constants, upload
    void UploadState() {
        gpu.allocate_arrays(prms, (int)hssoa.capacity);
        gpu.update_constants(prms);
        gpu.transfer_ponts_to_device(hssoa.host_buffer.data(), hssoa.capacity, hssoa.size);
    }
    // model_3d.cpp:132-137
    void Prepare() { gpu.update_constants(prms); }

    // model_3d.cpp:24-112; returns SimulationTime < SimulationEndTime
    bool Step() {
        gpu.reset_indenter_force_accumulator();
        const int N = prms.UpdateEveryNthStep;
        const int n = N - (prms.SimulationStep % N);
        const double t = gpu.step(n, prms.SimulationTime, prms.SimulationStep);
        prms.SimulationTime = t;
        prms.SimulationStep += n;
        gpu.transfer_from_device(hssoa.host_buffer.data(), hssoa.capacity, hssoa.size);
        const auto &F = gpu.indenter_force;
        indenter_force_log.push_back({t, std::sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]), F[0], F[1]});

        // model_3d.cpp:76-83, one partition: device capacity = host capacity
        const double n_pts = (double)(hssoa.size > 0 ? hssoa.size : 1);
        const double free_space = (double)(hssoa.capacity - hssoa.size) / n_pts;
        const double disabled = (double)hssoa.count_disabled() / (double)(hssoa.capacity > 0 ? hssoa.capacity : 1);
        if (free_space < prms.RebalanceThresholdFreeSpaceRemaining ||
            disabled > prms.RebalanceThresholdDisabledPercentage) {
            // model_3d.cpp:100-106
            gpu.remove_disabled_and_sort();
            gpu.transfer_from_device(hssoa.host_buffer.data(), hssoa.capacity, hssoa.size);
            rebalances++;
        }
        return prms.SimulationTime < prms.SimulationEndTime;
    }
};

} // namespace icempm

#endif
