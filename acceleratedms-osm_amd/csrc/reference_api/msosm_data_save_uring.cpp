// MSOSM_DataSave_Uring over the C ABI (data_save/msosm_data_save_uring.cpp:38-288).
#include "msosm_data_save_uring.h"

#include <vector>

MSOSM_DataSave_Uring::MSOSM_DataSave_Uring(float bw, float *dm, float f0, int numDMs)
    : MSOSM_GPU_BATCH(bw, dm, f0, numDMs)
{
}

MSOSM_DataSave_Uring::~MSOSM_DataSave_Uring()
{
    if (saver_) msosm_saver_destroy(saver_);
}

void MSOSM_DataSave_Uring::config_save(std::string dir, std::string prefix)
{
    // msosm_data_save_uring.cpp:40-71: names from the first two DMs' decimal places
    paths_.clear();
    std::vector<char> buf(4096);
    for (int i = 0; i < numDMs; ++i) {
        msosm::check(msosm_saver_file_name(dir.c_str(), prefix.c_str(), dm_values, numDMs, i, buf.data(), buf.size()),
                     "config_save");
        paths_.emplace_back(buf.data());
    }
}

void MSOSM_DataSave_Uring::initialize_uring(int slot_count, unsigned long save_count)
{
    // msosm_data_save_uring.cpp:73-101
    if (!plan_) msosm::check(MSOSM_ERR_STATE, "initialize_uring: call initialize_uint16 first");
    if (paths_.empty()) config_save();
    std::vector<const char *> p;
    for (auto &s : paths_) p.push_back(s.c_str());
    const size_t block = (size_t)batch * M * sizeof(uint16_pair);
    int flags = MSOSM_SAVE_PINNED | (pwrite_ ? 0 : MSOSM_SAVE_URING) | (direct_ ? MSOSM_SAVE_DIRECT : 0);
    if (saver_) msosm_saver_destroy(saver_);
    saver_ = nullptr;
    started_ = false;
    msosm::check(msosm_saver_create(p.data(), (int)p.size(), block, slot_count, save_count, flags, &saver_),
                 "initialize_uring");
}

void MSOSM_DataSave_Uring::copy_to_slot()
{
    // msosm_data_save_uring.cpp:103-113: wait for a free slot, D2H the output, mark it copying
    if (!saver_) msosm::check(MSOSM_ERR_STATE, "copy_to_slot: call initialize_uring first");
    void *slot = nullptr;
    msosm::check(msosm_saver_acquire(saver_, &slot), "copy_to_slot");
    msosm::check(msosm_plan_get_output(plan_, MSOSM_OUT_U16, slot), "copy_to_slot");
    msosm::check(msosm_saver_commit_stream(saver_, get_output_stream()), "copy_to_slot");
}

void MSOSM_DataSave_Uring::poll_cuda_state()
{
    // msosm_data_save_uring.cpp:115-154, on the calling thread (start_saving runs it on its own)
    if (!saver_) msosm::check(MSOSM_ERR_STATE, "poll_cuda_state: call initialize_uring first");
    started_ = true;
    msosm::check(msosm_saver_run_poller(saver_), "poll_cuda_state");
}

void MSOSM_DataSave_Uring::save_to_disk(int save_index)
{
    // :156-222: io_uring writes of one slot, on the calling thread
    if (!saver_) msosm::check(MSOSM_ERR_STATE, "save_to_disk: call initialize_uring first");
    started_ = true;
    msosm::check(msosm_saver_run_writer(saver_, save_index, 1), "save_to_disk");
}

void MSOSM_DataSave_Uring::save_to_disk_pwrite(int save_index)
{
    // :224-258: pwrite writes of one slot, on the calling thread
    if (!saver_) msosm::check(MSOSM_ERR_STATE, "save_to_disk_pwrite: call initialize_uring first");
    started_ = true;
    msosm::check(msosm_saver_run_writer(saver_, save_index, 0), "save_to_disk_pwrite");
}

void MSOSM_DataSave_Uring::start_saving()
{
    if (!saver_) msosm::check(MSOSM_ERR_STATE, "start_saving: call initialize_uring first");
    msosm::check(msosm_saver_start(saver_), "start_saving");
    started_ = true;
}

void MSOSM_DataSave_Uring::join_saving()
{
    // msosm_data_save_uring.cpp:270-288
    if (!saver_ || !started_) msosm::check(MSOSM_ERR_STATE, "join_saving: not saving");
    msosm::check(msosm_saver_join(saver_), "join_saving");
    synchronize();
}
