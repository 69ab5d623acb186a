// Runs output cycles of the C++ Model3D mirror (include/icempm_model3d.hpp)
// on a state written by the test that calls it, the way a host program built
// on the reference's Model3D drives its simulation loop.
//
//   model3d_cycles <in.bin> <out.bin> <n_cycles>
//
// in.bin:  icempm_SimParams3D, icempm_Indenter, int32 fuse, int32 sort_period,
//          uint32 capacity, uint32 size, capacity*28 doubles (HostSideSOA)
// out.bin: uint32 size, int32 rebalances, int32 cycles, cycles*4 doubles of
//          the indenter-force log, the last cycle's IndenterSubdivisions*GridZ*3
//          scaled sensor totals, capacity*28 doubles (HostSideSOA)
// Exit status 0 on success; a thrown std::runtime_error prints and exits 1.
#include <cstdint>
#include <cstdio>
#include <exception>
#include <fstream>
#include <string>

#include "icempm_model3d.hpp"

namespace {

template <typename T> void get(std::ifstream &f, T &v) {
    f.read(reinterpret_cast<char *>(&v), sizeof(T));
    if (!f) throw std::runtime_error("short input file");
}

template <typename T> void put(std::ofstream &f, const T &v) { f.write(reinterpret_cast<const char *>(&v), sizeof(T)); }

int run(const char *in_path, const char *out_path, int cycles) {
    icempm::Model3D model;
    icempm_Indenter ind{};
    int32_t fuse = 1, sort_period = 0;
    uint32_t capacity = 0, size = 0;
    std::ifstream in(in_path, std::ios::binary);
    if (!in) throw std::runtime_error(std::string("cannot open ") + in_path);
    get(in, model.prms);
    get(in, ind);
    get(in, fuse);
    get(in, sort_period);
    get(in, capacity);
    get(in, size);
    model.hssoa.Allocate(capacity);
    model.hssoa.size = size;
    in.read(reinterpret_cast<char *>(model.hssoa.host_buffer.data()),
            (std::streamsize)(model.hssoa.host_buffer.size() * sizeof(double)));
    if (!in) throw std::runtime_error("short point buffer");

    model.UploadState();
    model.gpu.set_indenter(ind);
    model.gpu.set_fusion(fuse != 0);
    model.gpu.set_sort_period(sort_period);
    for (int c = 0; c < cycles; c++) {
        model.Step();
        const auto &e = model.indenter_force_log.back();
        std::printf("cycle %d: t %.9g step %d |F| %.6e points %u rebalances %d\n", c, e[0], model.prms.SimulationStep,
                    e[1], model.hssoa.size, model.rebalances);
        std::fflush(stdout);
    }

    std::ofstream out(out_path, std::ios::binary);
    put(out, model.hssoa.size);
    put(out, (int32_t)model.rebalances);
    put(out, (int32_t)cycles);
    for (const auto &e : model.indenter_force_log)
        for (double v : e) put(out, v);
    for (double v : model.gpu.indenter_sensor_total) put(out, v);
    out.write(reinterpret_cast<const char *>(model.hssoa.host_buffer.data()),
              (std::streamsize)(model.hssoa.host_buffer.size() * sizeof(double)));
    if (!out) throw std::runtime_error(std::string("cannot write ") + out_path);
    return 0;
}

} // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <in.bin> <out.bin> <n_cycles>\n", argv[0]);
        return 2;
    }
    try {
        return run(argv[1], argv[2], std::stoi(argv[3]));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "model3d_cycles: %s\n", e.what());
        return 1;
    }
}
