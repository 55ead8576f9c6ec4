// MSOSM_DataSave_Uring: MS-OSM with uint16 output saved to one file per DM through a ring of pinned
// host slots (data_save/include/msosm_data_save_uring.h:22-52, data_save/msosm_data_save_uring.cpp),
// over the output pipeline of the C ABI (msosm_saver_*: io_uring by raw system calls).
#pragma once

#include <string>
#include <vector>

#include "msosm_gpu_batch.h"

// slot states of the reference (msosm_data_save_uring.h:13-20); the saver keeps them internally
enum class SaveSlotState { EMPTY, COPYING, READY, SAVING, DONE };

class MSOSM_DataSave_Uring : public MSOSM_GPU_BATCH {
public:
    MSOSM_DataSave_Uring(float bw, float *dm, float f0, int numDMs);
    ~MSOSM_DataSave_Uring() override;
    // files <dir>/<prefix>_DM<value>.dat, one per DM (the reference opens them here)
    void config_save(std::string dir = "", std::string prefix = "");
    // slot ring for save_count pushes; call after initialize_uint16 (save mode)
    void initialize_uring(int slot_count, unsigned long save_count);
    // queue the last push's uint16 output into the next free slot (waits while every slot is busy)
    void copy_to_slot();
    // the reference's thread bodies: here they run inside start_saving(); calling them is an error
    void poll_cuda_state();
    void save_to_disk(int save_index);
    void save_to_disk_pwrite(int save_index);
    void start_saving();
    void join_saving();

    // ---- extensions ----
    // pwrite writers instead of io_uring (the reference's save_to_disk_pwrite); before initialize_uring
    void use_pwrite(bool on) { pwrite_ = on; }
    // O_DIRECT files (default on, as the reference)
    void use_direct(bool on) { direct_ = on; }

private:
    std::vector<std::string> paths_;
    msosm_saver *saver_ = nullptr;
    bool pwrite_ = false, direct_ = true, started_ = false;
};
