/*
 * icempm_gpu_implementation6.hpp — header-only C++ mirror of the
 * reference's GPU_Implementation6 (simulation/gpu_implementation6.h:24-59)
 * for one partition, on top of the C-ABI in icempm_3dm.h.
 *
 * Member names and argument meaning follow the reference; every failing
 * call throws std::runtime_error with the library's message, as the
 * reference does (e.g. gpu_partition_3d.cu:67, gpu_implementation6.cpp:447).
 * The caller keeps owning its HostSideSOA buffer (host_side_soa.h:51-70):
 * pass host_buffer / capacity / size, as GPU_Partition_3D reads them.
 */
#ifndef ICEMPM_GPU_IMPLEMENTATION6_HPP
#define ICEMPM_GPU_IMPLEMENTATION6_HPP

#include <array>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "icempm_3dm.h"

namespace icempm {

inline void throw_if(int rc) {
    if (rc != ICEMPM_OK) throw std::runtime_error(std::string("icempm: ") + icempm_last_error());
}

class GPU_Implementation6 {
public:
    explicit GPU_Implementation6(int device = 0) : device_(device) {}
    ~GPU_Implementation6() { icempm_destroy(ctx_); }
    GPU_Implementation6(const GPU_Implementation6 &) = delete;
    GPU_Implementation6 &operator=(const GPU_Implementation6 &) = delete;

    // gpu_implementation6.cpp:319-337 (one partition)
    void initialize() {
        if (!ctx_) throw_if(icempm_create(device_, &ctx_));
    }
    // gpu_implementation6.cpp:385-401
    void allocate_arrays(const icempm_SimParams3D &prms, int points_capacity) {
        initialize();
        prms_ = prms;
        throw_if(icempm_allocate(ctx_, &prms, points_capacity));
        indenter_sensor_total.assign((size_t)3 * prms.IndenterSubdivisions * prms.GridZ, 0.0);
    }
    // gpu_implementation6.cpp:522-525
    void update_constants(const icempm_SimParams3D &prms) {
        prms_ = prms;
        throw_if(icempm_update_constants(ctx_, &prms));
    }
    void set_indenter(const icempm_Indenter &ind) { throw_if(icempm_set_indenter(ctx_, &ind)); }
    // extensions: damage model, hipGraph replay of the substep loop
    void set_damage(const icempm_Damage &dmg) { throw_if(icempm_set_damage(ctx_, &dmg)); }
    void set_graph(bool on) { throw_if(icempm_set_graph(ctx_, on ? 1 : 0)); }
    // gpu_implementation6.cpp:404-414
    void transfer_ponts_to_device(const double *host_buffer, unsigned capacity, unsigned size) {
        throw_if(icempm_transfer_points_to_device(ctx_, host_buffer, capacity, 0, (int)size));
    }
    void reset_grid() { throw_if(icempm_reset_grid(ctx_)); }
    void reset_indenter_force_accumulator() { throw_if(icempm_reset_indenter_force_accumulator(ctx_)); }
    void p2g() { throw_if(icempm_p2g(ctx_)); }
    void receive_halos() {} // one partition: no halos
    void update_nodes(double indenter_x, double indenter_y, double indenter_z = 0.0) {
        throw_if(icempm_update_nodes(ctx_, indenter_x, indenter_y, indenter_z));
    }
    void g2p(bool recordPQ, bool /*enablePointTransfer*/ = false) { throw_if(icempm_g2p(ctx_, recordPQ ? 1 : 0)); }
    // Model3D::Step substep loop (model_3d.cpp:35-51); returns the advanced time
    double step(int n_substeps, double simulation_time, int simulation_step) {
        throw_if(icempm_step(ctx_, n_substeps, &simulation_time, simulation_step));
        return simulation_time;
    }
    // gpu_implementation6.cpp:417-496: points, accumulators, error code,
    // cycle force, the 1/(N h^2) scaling of the sensor bins and the check
    // that the enabled points still number prms.nPtsTotal
    void transfer_from_device(double *host_buffer, unsigned capacity, unsigned &size) {
        int n = 0;
        std::vector<double> raw(indenter_sensor_total.size());
        throw_if(icempm_transfer_from_device(ctx_, host_buffer, capacity, 0, &n, raw.empty() ? nullptr : raw.data()));
        size = (unsigned)n;
        indenter_sensor_total = raw;
        evaluate_indenter_total_force();
        const double s = (double)prms_.UpdateEveryNthStep * prms_.cellsize * prms_.cellsize;
        for (double &v : indenter_sensor_total) v /= s;
        long long count = 0;
        for (unsigned i = 0; i < size; i++) {
            long long u;
            std::memcpy(&u, host_buffer + ICEMPM_IDX_UTILITY * (size_t)capacity + i, sizeof(u));
            if (!(u & ICEMPM_STATUS_DISABLED)) count++;
        }
        if (count != prms_.nPtsTotal)
            throw std::runtime_error("transfer_from_device(): active point count mismatch");
    }
    // gpu_implementation6.cpp:499-510
    void evaluate_indenter_total_force() {
        indenter_force = {0.0, 0.0, 0.0};
        for (size_t j = 0; j + 2 < indenter_sensor_total.size(); j += 3)
            for (int k = 0; k < 3; k++) indenter_force[k] += indenter_sensor_total[j + k];
        for (double &f : indenter_force) f /= (double)prms_.UpdateEveryNthStep;
    }
    // gpu_implementation6.cpp:513-520
    void synchronize() { throw_if(icempm_synchronize(ctx_)); }
    // extension: the indenter accumulators alone (no point copy); synchronizes
    void read_indenter_bins(std::vector<double> &bins) {
        bins.assign((size_t)prms_.IndenterSubdivisions * prms_.GridZ * 3, 0.0);
        throw_if(icempm_read_indenter_bins(ctx_, bins.data()));
    }
    // HostSideSOA::RemoveDisabledAndSort (host_side_soa.cpp:56-67), on the device
    int remove_disabled_and_sort() {
        int n = 0;
        throw_if(icempm_remove_disabled_and_sort(ctx_, &n));
        return n;
    }
    // extensions: fusion of G2P with the next P2G, brick re-sort period
    void set_fusion(bool on) { throw_if(icempm_set_fusion(ctx_, on ? 1 : 0)); }
    void set_sort_period(int period) { throw_if(icempm_set_sort_period(ctx_, period)); }

    std::array<double, 3> indenter_force{{0.0, 0.0, 0.0}};
    std::vector<double> indenter_sensor_total;

private:
    int device_;
    icempm_ctx *ctx_ = nullptr;
    icempm_SimParams3D prms_{};
};

} // namespace icempm

#endif
