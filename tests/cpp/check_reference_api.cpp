// Drives the reference-compatible C++ classes with the call sequences of the reference's harnesses
// and writes the results for tests/test_gpu_reference_api.py to compare with the oracle.
//
//   save  <bw> <f0> <batch> <dms.txt> <in.u16> <out.u16>       check_msosm.cpp:79-80,168-174
//   fold  <bw> <f0> <batch> <dm> <period> <nbins> <in.u16> <profile.txt>   msosm_psrdada.cpp:101-149
//   osm   <bw> <f0> <batch> <dm> <in.u16> <out.c64>            check_osm.cpp:66-67,127-133
//   uring <bw> <f0> <batch> <dms.txt> <in.u16> <dir> <uring|pwrite>   msosm_save_simulated.cpp:124-149
//   uneven <bw> <f0> <push_blocks> <dms.txt> <in.c64> <out.c64>   per-DM lengths (check_msosm.cpp pattern)
//   dada_info <file.dada> <n>                                   PSRDADA header and first samples (no GPU)
//   psrdada <file.dada> <dm> <period> <count> <nbins>           msosm_psrdada.cpp:88-153 (dual-pol fold)
//   simulate <bw> <dm> <f0> <period> <dummy> <valid> <out.u16>  check_msosm.cpp:128-133 (GPU simulator)
//   simulate_new <bw> <dm> <f0> <period> <repeat> <out.u16> <out.c64>   generate_pulsar_signal_new
//   simulate_pulse <bw> <dm> <f0> <period> <repeat> <pinned> <out.u16> <out.c64>   msosm_save_simulated.cpp:115
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "fold_gpu.h"
#include "msosm_data_save_uring.h"
#include "msosm_gpu_uneven.h"
#include "psrdada.h"
#include "msosm_gpu_batch.h"
#include "osm_gpu_batch.h"
#include "simulated_complex_signal.h"

static std::vector<uint16_pair> read_u16(const std::string &path)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    const size_t n = (size_t)f.tellg() / sizeof(uint16_pair);
    f.seekg(0);
    std::vector<uint16_pair> v(n);
    f.read(reinterpret_cast<char *>(v.data()), n * sizeof(uint16_pair));
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "dada_info") {
        PSRDADA file(argv[2]);
        file.initBuffer(std::stoul(argv[3]));
        file.readSamples();
        std::cout << "bw " << file.bw << " fc " << file.fc << " nbit " << file.nbit << " nchan " << file.nchan
                  << " ndim " << file.ndim << " npol " << file.npol << " readCount " << file.readCount << " out "
                  << file.outfileName << std::endl;
        for (int i = 0; i < 4; i++)
            std::cout << "s" << i << " " << file.pol1_in[i].first << " " << file.pol1_in[i].second << " "
                      << file.pol2_in[i].first << " " << file.pol2_in[i].second << std::endl;
        std::cout << "last " << file.pol1_in[2048].first << " " << file.pol2_in[2048].first << std::endl;
        return 0;
    }
    if (mode == "simulate_pulse") {
        SimulatedComplexSignal sim(std::stof(argv[2]), std::stof(argv[3]), std::stof(argv[4]), std::stof(argv[5]),
                                   "uint16");
        sim.generate_pulsar_signal(std::stoul(argv[6]), false, 0, std::stoi(argv[7]) != 0);
        std::ofstream of(argv[8], std::ios::binary);
        of.write(reinterpret_cast<const char *>(sim.signal_u16), sim.signal_size * sizeof(uint16_pair));
        std::ofstream oc(argv[9], std::ios::binary);
        oc.write(reinterpret_cast<const char *>(sim.signal), sim.signal_size * sizeof(fftwf_complex));
        std::cout << "Np " << sim.Np << " Nd " << sim.Nd << " range " << sim.range << " size " << sim.signal_size
                  << std::endl;
        return 0;
    }
    if (mode == "simulate_new") {
        SimulatedComplexSignal sim(std::stof(argv[2]), std::stof(argv[3]), std::stof(argv[4]), std::stof(argv[5]),
                                   "uint16");
        sim.generate_pulsar_signal_new(std::stoul(argv[6]));
        std::ofstream of(argv[7], std::ios::binary);
        of.write(reinterpret_cast<const char *>(sim.signal_u16), sim.signal_size * sizeof(uint16_pair));
        std::ofstream oc(argv[8], std::ios::binary);
        oc.write(reinterpret_cast<const char *>(sim.signal), sim.signal_size * sizeof(fftwf_complex));
        std::cout << "Np " << sim.Np << " Nd " << sim.Nd << " range " << sim.range << " size " << sim.signal_size
                  << std::endl;
        return 0;
    }
    if (mode == "simulate") {
        SimulatedComplexSignal sim(std::stof(argv[2]), std::stof(argv[3]), std::stof(argv[4]), std::stof(argv[5]),
                                   "uint16");
        sim.generate_pulsar_signal_block(std::stoul(argv[6]), std::stoul(argv[7]));
        std::ofstream of(argv[8], std::ios::binary);
        of.write(reinterpret_cast<const char *>(sim.signal_u16), sim.signal_size * sizeof(uint16_pair));
        std::cout << "Np " << sim.Np << " Nd " << sim.Nd << " range " << sim.range << " size " << sim.signal_size
                  << std::endl;
        return 0;
    }
    if (mode == "psrdada") {
        // run_psrdada_GPU_batch (msosm_psrdada.cpp:88-153): both polarisations dedispersed, amplitude
        // sqrt(|a|^2 + |b|^2) folded (one polarisation: power)
        const float dm = std::stof(argv[3]), period = std::stof(argv[4]);
        const int count = std::stoi(argv[5]);
        const unsigned long time_bin = std::stoul(argv[6]);
        PSRDADA file(argv[2]);
        const float bw = file.bw * 1e6f, f0 = (file.fc - file.bw / 2) * 1e6f;
        MSOSM_GPU_BATCH gpu1(bw, dm, f0);
        MSOSM_GPU_BATCH *gpu2 = file.npol == 2 ? new MSOSM_GPU_BATCH(bw, dm, f0) : nullptr;
        gpu1.initialize_uint16(0, count, true);
        if (gpu2) gpu2->initialize_uint16(0, count, true);
        const unsigned long M = gpu1.M;
        file.initBuffer(count * M);
        Fold_GPU fold(period, bw, count * M, file.outfileName, time_bin);
        bool first = true;
        unsigned long used = 0;
        for (unsigned long i = 0; i < file.readCount; i++) {
            file.readSamples();
            if (first) { // leading invalid data (code 0) is skipped (msosm_psrdada.cpp:131-137)
                if (file.pol1_in[0].first == 0) continue;
                first = false;
            }
            gpu1.filter_block_uint16(file.pol1_in);
            if (gpu2) {
                gpu2->filter_block_uint16(file.pol2_in);
                fold.calculate_intensity<MSOSM_GPU_BATCH>(&gpu1, gpu2);
            } else {
                fold.calculate_intensity<MSOSM_GPU_BATCH>(&gpu1);
            }
            fold.fold_data_phase();
            used++;
        }
        fold.get_folded_data();
        fold.write_to_file();
        delete gpu2;
        std::cout << "M " << M << " calls " << used << " of " << file.readCount << " out " << file.outfileName
                  << std::endl;
        return 0;
    }
    const float bw = std::stof(argv[2]), f0 = std::stof(argv[3]);
    const int batch = std::stoi(argv[4]);
    if (mode == "save") {
        std::vector<float> dms;
        std::ifstream df(argv[5]);
        for (float v; df >> v;) dms.push_back(v);
        auto input = read_u16(argv[6]);
        MSOSM_GPU_BATCH msosm(bw, dms.data(), f0, (int)dms.size());
        msosm.initialize_uint16(0, batch);
        const size_t process_len = (size_t)batch * msosm.M;
        const size_t count = input.size() / process_len;
        std::vector<uint16_pair> output(dms.size() * count * process_len);
        for (size_t k = 0; k < count; k++) {
            msosm.filter_block_uint16(input.data() + k * process_len);
            msosm.get_output(output.data() + k * process_len * dms.size());
            msosm.synchronize(); // pageable host buffers: finish each call before reusing them
        }
        msosm.synchronize();
        std::ofstream of(argv[7], std::ios::binary);
        of.write(reinterpret_cast<const char *>(output.data()), output.size() * sizeof(uint16_pair));
        std::cout << "M " << msosm.M << " delaycount " << msosm.delaycount << " calls " << count << std::endl;
    } else if (mode == "uneven") {
        std::vector<float> dms;
        std::ifstream df(argv[5]);
        for (float v; df >> v;) dms.push_back(v);
        std::ifstream inf(argv[6], std::ios::binary | std::ios::ate);
        const size_t n = (size_t)inf.tellg() / sizeof(Complex);
        inf.seekg(0);
        std::vector<Complex> input(n);
        inf.read(reinterpret_cast<char *>(input.data()), n * sizeof(Complex));
        MSOSM_GPU_UNEVEN msosm(bw, dms.data(), f0, (int)dms.size());
        msosm.initialize(batch, MSOSM_IN_C64, MSOSM_OUT_C64);
        const size_t per = (size_t)msosm.push_samples();
        const size_t count = n / per;
        std::vector<Complex> output(dms.size() * count * per);
        for (size_t k = 0; k < count; k++) {
            msosm.filter_block_c64(input.data() + k * per);
            for (size_t i = 0; i < dms.size(); i++) msosm.get_output((int)i, output.data() + (i * count + k) * per);
        }
        msosm.synchronize();
        std::ofstream of(argv[7], std::ios::binary);
        of.write(reinterpret_cast<const char *>(output.data()), output.size() * sizeof(Complex));
        std::cout << "classes " << msosm.num_classes() << " push " << per << " calls " << count << std::endl;
    } else if (mode == "uring") {
        std::vector<float> dms;
        std::ifstream df(argv[5]);
        for (float v; df >> v;) dms.push_back(v);
        auto input = read_u16(argv[6]);
        MSOSM_DataSave_Uring msosm(bw, dms.data(), f0, (int)dms.size());
        msosm.use_pwrite(std::string(argv[8]) == "pwrite");
        msosm.initialize_uint16(0, batch);
        const size_t process_len = (size_t)batch * msosm.M;
        const int count = (int)(input.size() / process_len);
        msosm.config_save(argv[7], "simulated");
        msosm.initialize_uring(4, count);
        msosm.start_saving();
        for (int k = 0; k < count; k++) {
            msosm.filter_block_uint16(input.data() + k * process_len);
            msosm.copy_to_slot();
        }
        msosm.join_saving();
        std::cout << "M " << msosm.M << " calls " << count << std::endl;
    } else if (mode == "fold") {
        const float dm = std::stof(argv[5]), period = std::stof(argv[6]);
        const unsigned long nbins = std::stoul(argv[7]);
        auto input = read_u16(argv[8]);
        MSOSM_GPU_BATCH msosm(bw, dm, f0);
        msosm.initialize_uint16(0, batch, true);
        const size_t process_len = (size_t)batch * msosm.M;
        Fold_GPU fold(period, bw, process_len, argv[9], nbins);
        for (size_t k = 0; k < input.size() / process_len; k++) {
            msosm.filter_block_uint16(input.data() + k * process_len);
            fold.calculate_intensity<MSOSM_GPU_BATCH>(&msosm);
            fold.fold_data_phase();
        }
        fold.get_folded_data();
        fold.write_to_file();
        std::cout << "M " << msosm.M << " bins " << fold.time_bin << std::endl;
    } else if (mode == "osm") {
        const float dm = std::stof(argv[5]);
        auto input = read_u16(argv[6]);
        OSM_GPU_BATCH osm(bw, dm, f0);
        osm.initialize_uint16(0, batch);
        const size_t process_len = (size_t)batch * osm.M;
        const size_t count = input.size() / process_len;
        std::vector<Complex> output(count * process_len);
        for (size_t k = 0; k < count; k++) {
            osm.filter_block_uint16(input.data() + k * process_len);
            osm.get_output(output.data() + k * process_len);
        }
        std::ofstream of(argv[7], std::ios::binary);
        of.write(reinterpret_cast<const char *>(output.data()), output.size() * sizeof(Complex));
        std::cout << "M " << osm.M << " calls " << count << std::endl;
    } else {
        return 2;
    }
    return 0;
}
