// Internal plan hooks shared by the C-ABI files compiled with hipcc: the uneven-grid driver runs the
// forward part of several plans' pushes, then one dedispersion launch over all of them.
#pragma once

#include "kernels.h"
#include "msosm_amd.h"

namespace msosm {

struct HistArgs {
    float2 *hist;
    const void *in;
    int fmt;
    long long first;
    int M;
};
struct BinsArgs {
    uint16_t *bins;
    long long n;
    unsigned long nbins;
    double t_start, dt, period;
    int q = 1;       // > 1: in the order of a sub-unit accumulate launch of q sub-units (launch_bins_sub)
    long long m = 0; //   samples per block
};
// every launch argument of one push
struct PushArgs {
    int log2L = 0, fwd_grid = 0;
    FwdArgs fwd{};
    HistArgs hist{};
    bool has_bins = false;
    BinsArgs bins{};
    InvArgs inv{};
};
// allow_sub: the caller runs the plan's own dedispersion launch, which may be the sub-unit accumulate
// form (the uneven grid's one-launch kernel and tangent pushes read the natural layouts)
void plan_push_args(msosm_plan *p, const void *src_device, PushArgs *out, bool allow_sub);
// forward transforms, overlap history and fold bins of one push; fills the dedispersion arguments
int plan_push_prepare(msosm_plan *p, const void *src_device, InvArgs *out, bool allow_sub);
// advance the ring position and the fold clock after the dedispersion launch
void plan_push_commit(msosm_plan *p);
// MSOSM_OK when every DM has a design and an output or fold is selected (sets the error otherwise)
int plan_check_ready(const msosm_plan *p);
int plan_log2L(const msosm_plan *p);
int plan_workgroups(const msosm_plan *p);
int plan_num_units(const msosm_plan *p); // DMs * blocks
bool plan_ring32(const msosm_plan *p);
int plan_launch_mode(const msosm_plan *p);
hipStream_t plan_stream(const msosm_plan *p);

} // namespace msosm
