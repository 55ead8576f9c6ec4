// pybind11 module msosm_amd._core: a folded-profile engine for DM refinement over the C ABI.
//
// A ProfileEngine owns one plan for a batch of DM trials, plus a primal and a tangent fold. M and
// the ring depth (delaycount) are fixed at creation for DMs up to dm_max; every set_dms call designs
// each trial anew at its own DM (delay blocks and H), so the profile of a trial is this library's
// MS-OSM at exactly that DM (the float DM of the reference interface). For the current DMs
// it computes the folded profile of a device-resident record (forward), or the profile together
// with its derivative with respect to each DM (tangent: every block is dedispersed with H and
// with dH/dDM, and 2 Re(conj(y) dy/dDM) is folded with the primal bins). Nothing is stored between
// calls besides the profiles: the tangent pass recomputes the series.
//
// The design follows Prepare_MSOSM (dedispersion/msosm/msosm_prepare.cpp:3-69,215-284) for M and
// the segmentation, and generate_dedisp_params (:107-164) for H; the fold is Fold_GPU's
// (folding/fold_gpu.cpp:48-60) with one profile per DM.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "msosm_amd.h"

namespace py = pybind11;

namespace {

void check(int rc, const char *what)
{
    if (rc != MSOSM_OK) throw std::runtime_error(std::string(what) + ": " + msosm_last_error());
}

class ProfileEngine {
  public:
    ProfileEngine(float bw, float f0, py::array_t<float, py::array::c_style | py::array::forcecast> start_dms,
                  int batch, unsigned long nbins, double period, int in_format, int fftpoint, int phase_mode,
                  int device, float dm_max)
        : bw_(bw), f0_(f0), B_(batch), nbins_(nbins), period_(period), in_format_(in_format),
          phase_mode_(phase_mode)
    {
        const py::ssize_t D = start_dms.size();
        if (D <= 0) throw std::invalid_argument("ProfileEngine: need at least one DM");
        D_ = (int)D;
        start_.assign(start_dms.data(), start_dms.data() + D);
        dm_max_ = std::max(dm_max, *std::max_element(start_.begin(), start_.end()));
        // common M: the largest minimum order over the DMs up to dm_max (msosm_prepare.cpp:48-69), or
        // fftpoint/2 when that is not smaller (msosm_gpu_batch.cpp:32-44)
        std::vector<float> sizing(start_);
        sizing.push_back(dm_max_);
        int M = 0;
        for (float dm : sizing) {
            int nd = 0, m = 0;
            check(msosm_design_nd(bw, f0, dm, &nd), "design_nd");
            if (nd > 0) check(msosm_design_min_order(nd, &m), "design_min_order");
            M = std::max(M, m);
        }
        if (M == 0) M = 1 << 12;
        if (fftpoint > 0 && fftpoint / 2 >= M) M = fftpoint / 2;
        M_ = M;
        delays_.resize((size_t)D_ * 2 * M_);
        // ring depth for the largest DM allowed (the delaycount grows with the DM at a fixed M)
        int dc = 1;
        for (float dm : sizing) {
            int dci = 0;
            check(msosm_design_segmentation(bw, f0, dm, M_, delays_.data(), &dci), "design_segmentation");
            dc = std::max(dc, dci);
        }
        dc_ = dc;
        msosm_plan_desc desc{};
        desc.M = M_;
        desc.num_dms = D_;
        desc.batch = B_;
        desc.delaycount = dc_;
        desc.in_format = in_format;
        desc.out_flags = MSOSM_OUT_NONE;
        desc.device = device;
        desc.stream = nullptr;
        check(msosm_plan_create(&desc, &plan_), "plan_create");
        const unsigned long size = (unsigned long)B_ * M_;
        check(msosm_fold_create(period, bw, size, nbins, D_, device, &fold_), "fold_create");
        check(msosm_fold_create(period, bw, size, nbins, D_, device, &tfold_), "fold_create");
        check(msosm_plan_attach_fold(plan_, fold_), "plan_attach_fold");
        set_dms(start_dms, false);
    }

    ~ProfileEngine()
    {
        if (plan_) msosm_plan_destroy(plan_);
        if (fold_) msosm_fold_destroy(fold_);
        if (tfold_) msosm_fold_destroy(tfold_);
    }

    // new trial DMs: each trial's delay blocks (msosm_prepare.cpp:215-284) and H are designed at its
    // own DM; with_tangent also designs and uploads dH/dDM, which the tangent run needs. The delay
    // blocks are piecewise constant in DM, so dH/dDM is the derivative of the profile between the DMs
    // where a bin changes block.
    void set_dms(py::array_t<float, py::array::c_style | py::array::forcecast> dms, bool with_tangent)
    {
        if (dms.size() != D_) throw std::invalid_argument("set_dms: expected " + std::to_string(D_) + " DMs");
        cur_.assign(dms.data(), dms.data() + D_);
        for (int i = 0; i < D_; ++i)
            if (!(cur_[i] >= 0.0f) || cur_[i] > dm_max_)
                throw std::invalid_argument("set_dms: DM " + std::to_string(cur_[i]) + " outside [0, dm_max=" +
                                            std::to_string(dm_max_) + "] of this engine");
        std::vector<float> H((size_t)D_ * 4 * M_);
        {
            py::gil_scoped_release nogil;
            for (int i = 0; i < D_; ++i) {
                int dci = 0;
                check(msosm_design_segmentation(bw_, f0_, cur_[i], M_, delays_.data() + (size_t)i * 2 * M_, &dci),
                      "design_segmentation");
            }
            check(msosm_design_transfer_functions(bw_, f0_, cur_.data(), D_, M_, phase_mode_, H.data(), 0),
                  "design_transfer_functions");
            for (int i = 0; i < D_; ++i)
                check(msosm_plan_set_dm(plan_, i, delays_.data() + (size_t)i * 2 * M_, H.data() + (size_t)i * 4 * M_),
                      "plan_set_dm");
            tangent_ready_ = false;
            if (with_tangent) {
                // dH/dDM per DM on all host cores (H's buffer is reused)
                const int nt = std::max(1, std::min<int>(D_, (int)std::thread::hardware_concurrency()));
                std::vector<int> rc(nt, MSOSM_OK);
                std::vector<std::thread> pool;
                for (int t = 0; t < nt; ++t)
                    pool.emplace_back([&, t] {
                        for (int i = t; i < D_ && rc[t] == MSOSM_OK; i += nt)
                            rc[t] = msosm_design_transfer_function_ddm(bw_, f0_, cur_[i], M_, phase_mode_,
                                                                       H.data() + (size_t)i * 4 * M_);
                    });
                for (auto &th : pool) th.join();
                for (int r : rc) check(r, "design_transfer_function_ddm");
                for (int i = 0; i < D_; ++i)
                    check(msosm_plan_set_dm_tangent(plan_, i, H.data() + (size_t)i * 4 * M_), "plan_set_dm_tangent");
                tangent_ready_ = true;
            }
        }
    }

    // Fold a device-resident record of n_samples (a multiple of batch*M) from zero history.
    // Returns (profile [D][nbins] float64, d profile / d DM or None, counts [D][nbins] uint64).
    py::tuple run(std::uintptr_t input, long long n_samples, bool tangent)
    {
        const long long per = (long long)B_ * M_;
        if (n_samples <= 0 || n_samples % per)
            throw std::invalid_argument("run: n_samples must be a positive multiple of batch*M = " + std::to_string(per));
        if (tangent && !tangent_ready_) throw std::runtime_error("run: tangent needs set_dms(..., with_tangent=True)");
        const size_t bytes_per = in_format_ == MSOSM_IN_U16 ? 4 : in_format_ == MSOSM_IN_I8 ? 2 : 8;
        const size_t nb = (size_t)D_ * nbins_;
        std::vector<double> sums(nb), tsums(nb);
        std::vector<unsigned long long> cnts(nb), tcnts(nb);
        {
            py::gil_scoped_release nogil;
            check(msosm_plan_reset(plan_), "plan_reset");
            check(msosm_fold_reset(fold_), "fold_reset");
            check(msosm_fold_reset(tfold_), "fold_reset");
            const char *base = reinterpret_cast<const char *>(input);
            for (long long s = 0; s < n_samples; s += per) {
                const void *src = base + (size_t)s * bytes_per;
                if (tangent)
                    check(msosm_plan_push_tangent(plan_, src, MSOSM_INPUT_DEVICE, tfold_), "plan_push_tangent");
                else
                    check(msosm_plan_push(plan_, src, MSOSM_INPUT_DEVICE), "plan_push");
            }
            check(msosm_plan_synchronize(plan_), "plan_synchronize");
            check(msosm_fold_read(fold_, nullptr, sums.data(), cnts.data()), "fold_read");
            if (tangent) check(msosm_fold_read(tfold_, nullptr, tsums.data(), tcnts.data()), "fold_read");
        }
        py::array_t<double> prof({(py::ssize_t)D_, (py::ssize_t)nbins_});
        py::array_t<unsigned long long> counts({(py::ssize_t)D_, (py::ssize_t)nbins_});
        auto p = prof.mutable_data();
        auto c = counts.mutable_data();
        for (size_t i = 0; i < nb; ++i) {
            p[i] = cnts[i] ? sums[i] / (double)cnts[i] : 0.0;
            c[i] = cnts[i];
        }
        if (!tangent) return py::make_tuple(prof, py::none(), counts);
        for (size_t i = 0; i < nb; ++i)
            if (tcnts[i] != cnts[i]) throw std::runtime_error("run: tangent and primal fold counts differ");
        py::array_t<double> dprof({(py::ssize_t)D_, (py::ssize_t)nbins_});
        auto dp = dprof.mutable_data();
        for (size_t i = 0; i < nb; ++i) dp[i] = cnts[i] ? tsums[i] / (double)cnts[i] : 0.0;
        return py::make_tuple(prof, dprof, counts);
    }

    int M() const { return M_; }
    int delaycount() const { return dc_; }
    int num_dms() const { return D_; }
    int batch() const { return B_; }
    unsigned long nbins() const { return nbins_; }
    py::array_t<float> dms() const { return py::array_t<float>((py::ssize_t)D_, cur_.data()); }
    py::array_t<float> start_dms() const { return py::array_t<float>((py::ssize_t)D_, start_.data()); }
    float dm_max() const { return dm_max_; }
    // the delay block of every bin of trial i at its current DM (int32 [2M])
    py::array_t<int> delays(int i) const
    {
        if (i < 0 || i >= D_) throw std::out_of_range("delays: trial index");
        return py::array_t<int>((py::ssize_t)2 * M_, delays_.data() + (size_t)i * 2 * M_);
    }

  private:
    float bw_, f0_;
    int B_;
    unsigned long nbins_;
    double period_;
    int in_format_, phase_mode_;
    int D_ = 0, M_ = 0, dc_ = 0;
    float dm_max_ = 0.0f;
    bool tangent_ready_ = false;
    std::vector<float> start_, cur_;
    std::vector<int> delays_;
    msosm_plan *plan_ = nullptr;
    msosm_fold *fold_ = nullptr, *tfold_ = nullptr;
};

} // namespace

PYBIND11_MODULE(_core, m)
{
    m.doc() = "Folded-profile engine with DM derivatives (MS-OSM dedispersion + fold on the GPU)";
    py::class_<ProfileEngine>(m, "ProfileEngine")
        .def(py::init<float, float, py::array_t<float, py::array::c_style | py::array::forcecast>, int, unsigned long,
                      double, int, int, int, int, float>(),
             py::arg("bw"), py::arg("f0"), py::arg("start_dms"), py::arg("batch"), py::arg("nbins"),
             py::arg("period"), py::arg("in_format") = (int)MSOSM_IN_C64, py::arg("fftpoint") = 0,
             py::arg("phase_mode") = (int)MSOSM_PHASE_FP64, py::arg("device") = 0, py::arg("dm_max") = 0.0f)
        .def("set_dms", &ProfileEngine::set_dms, py::arg("dms"), py::arg("with_tangent") = false)
        .def("run", &ProfileEngine::run, py::arg("input"), py::arg("n_samples"), py::arg("tangent") = false)
        .def_property_readonly("M", &ProfileEngine::M)
        .def_property_readonly("delaycount", &ProfileEngine::delaycount)
        .def_property_readonly("num_dms", &ProfileEngine::num_dms)
        .def_property_readonly("batch", &ProfileEngine::batch)
        .def_property_readonly("nbins", &ProfileEngine::nbins)
        .def_property_readonly("dms", &ProfileEngine::dms)
        .def_property_readonly("start_dms", &ProfileEngine::start_dms)
        .def_property_readonly("dm_max", &ProfileEngine::dm_max)
        .def("delays", &ProfileEngine::delays, py::arg("i"));
}
