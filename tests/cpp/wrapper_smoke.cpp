// Compile check of the C++ mirror of GPU_Implementation6: a caller written
// the way Model3D::Step (model_3d.cpp:24-112) drives the reference class.
#include "icempm_gpu_implementation6.hpp"

double run_cycle(icempm::GPU_Implementation6 &gpu, icempm_SimParams3D &prms, double *buf, unsigned capacity,
                 unsigned &size) {
    gpu.reset_indenter_force_accumulator();
    double t = prms.SimulationTime;
    int count = 0;
    do {
        t += prms.InitialTimeStep;
        prms.indenter_x = prms.indenter_x_initial + t * prms.IndVelocity;
        gpu.reset_grid();
        gpu.p2g();
        gpu.receive_halos();
        gpu.update_nodes(prms.indenter_x, prms.indenter_y);
        const bool isZeroStep = (prms.SimulationStep + count) % prms.UpdateEveryNthStep == 0;
        gpu.g2p(isZeroStep, false);
        count++;
    } while ((prms.SimulationStep + count) % prms.UpdateEveryNthStep != 0);
    prms.SimulationTime = t;
    prms.SimulationStep += count;
    gpu.transfer_from_device(buf, capacity, size);
    return gpu.indenter_force[0];
}
