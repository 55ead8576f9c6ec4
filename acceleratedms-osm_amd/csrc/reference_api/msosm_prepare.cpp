// Prepare_MSOSM (dedispersion/msosm/msosm_prepare.cpp:3-284) over the design entry points of the C ABI.
#include "msosm_prepare.h"

#include <algorithm>

Prepare_MSOSM::Prepare_MSOSM(float bw, float dm, float f0) : numDMs(1), bw(bw), f0(f0)
{
    // msosm_prepare.cpp:3-23
    fs = bw;
    kdm = (float)(4.15 * 1e15);
    w0 = (float)(2 * pi * f0);
    dm_values = new float[1]{dm};
    own_dm_values = true;
    Nd_values = new int[1]{0};
    if (bw > 0 && f0 > 0) msosm::check(msosm_design_nd(bw, f0, dm, &Nd_values[0]), "Prepare_MSOSM");
    start_Nd = end_Nd = Nd_values[0];
    if (verbose) {
        cout << "Sampling frequency: " << fs << endl;
        cout << "Start frequency: " << f0 << endl;
        cout << "Dispersion smearing points: " << Nd_values[0] << endl;
    }
    dedisp_params = new fftwf_complex *[numDMs]();
    delay_points = new int *[numDMs]();
    delaycount_values = new int[numDMs]();
}

Prepare_MSOSM::Prepare_MSOSM(float bw, float *dm, float f0, int numDMs) : numDMs(numDMs), bw(bw), dm_values(dm), f0(f0)
{
    // msosm_prepare.cpp:25-46
    fs = bw;
    kdm = (float)(4.15 * 1e15);
    w0 = (float)(2 * pi * f0);
    Nd_values = new int[numDMs]();
    for (int i = 0; i < numDMs; i++) {
        msosm::check(msosm_design_nd(bw, f0, dm_values[i], &Nd_values[i]), "Prepare_MSOSM");
        if (verbose) {
            cout << "Sampling frequency: " << fs << endl;
            cout << "Start frequency: " << f0 << endl;
            cout << "Dispersion smearing points for DM " << dm_values[i] << ": " << Nd_values[i] << endl;
        }
    }
    start_Nd = Nd_values[0];
    end_Nd = Nd_values[numDMs - 1];
    dedisp_params = new fftwf_complex *[numDMs]();
    delay_points = new int *[numDMs]();
    delaycount_values = new int[numDMs]();
}

void Prepare_MSOSM::release_tables()
{
    for (int i = 0; i < numDMs; i++) {
        delete[] dedisp_params[i];
        dedisp_params[i] = nullptr;
        delete[] delay_points[i];
        delay_points[i] = nullptr;
    }
}

Prepare_MSOSM::~Prepare_MSOSM()
{
    release_tables();
    delete[] dedisp_params;
    delete[] delay_points;
    delete[] delaycount_values;
    delete[] Nd_values;
    delete[] M_values;
    if (own_dm_values) delete[] dm_values;
}

void Prepare_MSOSM::calculate_min_order()
{
    // msosm_prepare.cpp:48-69
    if (verbose) cout << "Prepare parameters..." << endl;
    delete[] M_values;
    M_values = new int[numDMs];
    for (int i = 0; i < numDMs; i++) {
        msosm::check(msosm_design_min_order(Nd_values[i], &M_values[i]), "calculate_min_order");
        if (verbose) cout << "Minimum filter order for DM " << dm_values[i] << ": " << M_values[i] << endl;
    }
    M = *std::max_element(M_values, M_values + numDMs);
    M_min = M;
}

void Prepare_MSOSM::override_order(int order) { M = order; }

void Prepare_MSOSM::generate_dedisp_params(int i)
{
    // msosm_prepare.cpp:107-164: H of DM i, 2M complex values
    delete[] dedisp_params[i];
    dedisp_params[i] = new fftwf_complex[2 * (size_t)M];
    msosm::check(msosm_design_transfer_function(bw, f0, dm_values[i], M, phase_mode,
                                                reinterpret_cast<float *>(dedisp_params[i])),
                 "generate_dedisp_params");
}

void Prepare_MSOSM::segmentation(int i)
{
    // msosm_prepare.cpp:215-284
    delete[] delay_points[i];
    delay_points[i] = new int[2 * (size_t)M];
    int dc = 0;
    msosm::check(msosm_design_segmentation(bw, f0, dm_values[i], M, delay_points[i], &dc), "segmentation");
    delaycount_values[i] = dc;
    if (dc > delaycount) delaycount = dc;
    if (verbose) {
        cout << "Delaycount for DM " << dm_values[i] << ": " << delaycount_values[i] << endl;
        line();
    }
}
