// PyTorch extension with the reference's binding names (src/maps/map.cpp:16-44),
// forwarding to the C ABI of libeggsync.so (include/egg_sync.h).  No HIP
// headers: a device tensor's stream comes in from Python as an integer handle.
//
//   EGG_SynC_wrapper(D, eps)                      map.cpp:31-33 -> EGG_SynC.cu:1229
//   DynamicalClustering_wrapper(D, eps, lam)      map.cpp:16-18 -> SynC.cpp:268 (run on the GPU)
//   DynamicalClustering_parallel_wrapper(...)     map.cpp:20-22 -> SynC.cpp:328 (same GPU path)
//   GPU_DynamicalClustering_wrapper(D, eps, lam)  map.cpp:27-29 -> simple_GPU_SynC.cu:105
//                                                 (brute force, labels float32)
// Each returns [labels, itr_times, stage_times, space] like the reference
// (EGG_SynC.cu:1587-1600, SynC.cpp:309-325): labels float32 for EGG-SynC and
// int32 for SynC, itr_times float32 seconds per iteration, stage_times
// float32 seconds for the six stages, space int64 device bytes.
#include <torch/extension.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "egg_sync.h"

namespace {

void check_input(const at::Tensor &D) {
    TORCH_CHECK(D.dim() == 2, "D must be a 2-D tensor [n, d]");
    TORCH_CHECK(D.scalar_type() == at::kFloat, "D must be float32");
    TORCH_CHECK(D.size(1) >= 1 && D.size(1) <= EGG_SYNC_MAX_D, "d must be in [1, ",
                EGG_SYNC_MAX_D, "]");
}

void raise_if_error(int rc) {
    if (rc < 0) throw std::runtime_error(std::string("egg_sync: ") + egg_sync_error_string(rc));
}

std::vector<at::Tensor> run(at::Tensor D, float eps, int stop_rule, int label_rule, float lam,
                            bool labels_float, int64_t stream, int max_iter) {
    check_input(D);
    D = D.contiguous();
    const int64_t n = D.size(0);
    const int32_t d = (int32_t)D.size(1);
    egg_sync_options opt;
    egg_sync_default_options(&opt);
    opt.stop_rule = stop_rule;
    opt.label_rule = label_rule;
    opt.lam = lam;
    opt.timing = 1;
    opt.max_iter = max_iter;
    std::vector<float> itr(16384, 0.f);
    egg_sync_stats st{};
    st.itr_ms = itr.data();
    st.itr_ms_cap = (int32_t)itr.size();
    at::Tensor labels;
    int rc;
    if (D.is_cuda()) {
        labels = at::empty({n}, D.options().dtype(at::kInt));
        pybind11::gil_scoped_release nogil;
        rc = egg_sync_solve_device(D.data_ptr<float>(), n, d, eps, &opt, labels.data_ptr<int>(),
                                   nullptr, &st, reinterpret_cast<void *>(stream));
    } else {
        labels = at::empty({n}, at::kInt);
        pybind11::gil_scoped_release nogil;
        rc = egg_sync_solve_ex(D.data_ptr<float>(), n, d, eps, &opt, labels.data_ptr<int>(),
                               nullptr, &st);
    }
    raise_if_error(rc);
    const int iters = st.iterations;
    at::Tensor itr_times = at::empty({iters}, at::kFloat);
    for (int i = 0; i < iters; i++)
        itr_times[i] = i < st.itr_ms_cap ? itr[i] / 1000.f : 0.f;
    at::Tensor stage_times = at::empty({6}, at::kFloat);
    for (int k = 0; k < 6; k++) stage_times[k] = (float)(st.stage_ms[k] / 1000.0);
    at::Tensor space = at::full({1}, st.workspace_bytes, at::kLong);
    at::Tensor lab = labels_float ? labels.to(at::kFloat) : labels;
    return {lab, itr_times, stage_times, space};
}

std::vector<at::Tensor> EGG_SynC_wrapper(at::Tensor &D, float eps, int64_t stream) {
    return run(D, eps, EGG_SYNC_STOP_EGG, EGG_SYNC_LABELS_CELLS, 1.0f - 1e-3f, true, stream, 0);
}

std::vector<at::Tensor> DynamicalClustering_wrapper(at::Tensor &D, float eps, float lam,
                                                    int64_t stream) {
    return run(D, eps, EGG_SYNC_STOP_RC, EGG_SYNC_LABELS_SYNC, lam, false, stream, 0);
}

// FSynC (SynC.cpp:426-487, map.cpp:21-23): the SynC iteration with R-tree range
// queries (node capacity B) instead of a scan over all points.  The neighbour
// sets are the same, so on the GPU it is the SynC solve above; B shaped only
// the reference's CPU tree and is checked, not used.
std::vector<at::Tensor> FDynamicalClustering_wrapper(at::Tensor &D, float eps, float lam,
                                                     int64_t B, int64_t stream) {
    if (B < 2) throw std::invalid_argument("FDynamicalClustering: node capacity B must be >= 2");
    return DynamicalClustering_wrapper(D, eps, lam, stream);
}

std::vector<at::Tensor> GPU_DynamicalClustering_wrapper(at::Tensor &D, float eps, float lam,
                                                        int64_t stream) {
    check_input(D);
    D = D.contiguous();
    const int64_t n = D.size(0);
    const int32_t d = (int32_t)D.size(1);
    std::vector<float> itr(4096, 0.f);
    egg_sync_stats st{};
    st.itr_ms = itr.data();
    st.itr_ms_cap = (int32_t)itr.size();
    at::Tensor labels;
    int rc;
    if (D.is_cuda()) {
        labels = at::empty({n}, D.options().dtype(at::kInt));
        pybind11::gil_scoped_release nogil;
        rc = egg_sync_simple_solve_device(D.data_ptr<float>(), n, d, eps, lam, 0,
                                          labels.data_ptr<int>(), nullptr, &st,
                                          reinterpret_cast<void *>(stream));
    } else {
        labels = at::empty({n}, at::kInt);
        pybind11::gil_scoped_release nogil;
        rc = egg_sync_simple_solve(D.data_ptr<float>(), n, d, eps, lam, 0, labels.data_ptr<int>(),
                                   nullptr, &st);
    }
    raise_if_error(rc);
    const int iters = st.iterations;
    at::Tensor itr_times = at::empty({iters}, at::kFloat);
    for (int i = 0; i < iters; i++)
        itr_times[i] = i < st.itr_ms_cap ? itr[i] / 1000.f : 0.f;
    at::Tensor stage_times = at::empty({6}, at::kFloat);
    for (int k = 0; k < 6; k++) stage_times[k] = (float)(st.stage_ms[k] / 1000.0);
    at::Tensor space = at::full({1}, st.workspace_bytes, at::kLong);
    return {labels.to(at::kFloat), itr_times, stage_times, space};
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("EGG_SynC_wrapper", &EGG_SynC_wrapper, "EGG-SynC solve (EGG_SynC.cu:1229)",
          pybind11::arg("D"), pybind11::arg("eps"), pybind11::arg("stream") = 0);
    m.def("DynamicalClustering_wrapper", &DynamicalClustering_wrapper,
          "SynC solve, r_c stop (SynC.cpp:268), on the GPU", pybind11::arg("D"),
          pybind11::arg("eps"), pybind11::arg("lam"), pybind11::arg("stream") = 0);
    m.def("GPU_DynamicalClustering_wrapper", &GPU_DynamicalClustering_wrapper,
          "brute-force SynC on the GPU (simple_GPU_SynC.cu:105)", pybind11::arg("D"),
          pybind11::arg("eps"), pybind11::arg("lam"), pybind11::arg("stream") = 0);
    m.def("DynamicalClustering_parallel_wrapper", &DynamicalClustering_wrapper,
          "SynC solve, r_c stop (SynC.cpp:328), on the GPU", pybind11::arg("D"),
          pybind11::arg("eps"), pybind11::arg("lam"), pybind11::arg("stream") = 0);
    m.def("FDynamicalClustering_wrapper", &FDynamicalClustering_wrapper,
          "FSynC solve (SynC.cpp:426), the SynC solve on the GPU", pybind11::arg("D"),
          pybind11::arg("eps"), pybind11::arg("lam"), pybind11::arg("B"),
          pybind11::arg("stream") = 0);
}
